// Command-line front end of the engine for t-SNE: the t-SNE subset of
// tapkee's CLI (src/cli/main.cpp:158-548, src/cli/util.hpp:73-136) plus a
// --seed option (the reference seeds std::rand with time(NULL),
// main.cpp:160, so its runs cannot be repeated).
//
//   tapkee -i data.csv -o embedding.csv -m t-sne --sne-perplexity 30 --sne-theta 0.5 --seed 1
//
// Input: delimiter-separated real values, one vector per line (one per
// column with --transpose-input).  Output: the embedding, one vector per
// line (per column with --transpose-output), written with the stream's
// default formatting as the reference's write_matrix (util.hpp:124-136).
// Options of other methods are accepted and checked as the reference
// checks them; methods other than t-SNE are reported as not provided.
// Differences: the default method is t-sne (the reference's is
// locally_linear_embedding, which this engine does not provide), and the
// neighbour search is always the exact GPU one whatever --neighbors-method
// says.
#include <tapkee_amd/tapkee.hpp>

#include <algorithm>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Option {
    std::string lng, shrt, def, help;
    bool flag;
};

const std::vector<Option> kOptions = {
    {"input-file", "i", "/dev/stdin", "Input filename (delimiter-separated matrix of real values)", false},
    {"transpose-input", "", "", "A column of the input is one vector (default: a line is one vector)", true},
    {"transpose-output", "", "", "Write one embedding vector per column (default: per line)", true},
    {"output-file", "o", "/dev/stdout", "Output filename", false},
    {"output-projection-matrix-file", "opmat", "/dev/null", "Projection matrix file (t-SNE has none)", false},
    {"output-projection-mean-file", "opmean", "/dev/null", "Projection mean file (t-SNE has none)", false},
    {"delimiter", "d", ",", "Delimiter to be used in reading and writing matrices", false},
    {"help", "h", "", "Print usage of the program", true},
    {"benchmark", "", "", "Output benchmarking information about the time of algorithm steps", true},
    {"verbose", "", "", "Be more verbose in logging", true},
    {"debug", "", "", "Output debugging information", true},
    {"method", "m", "t-sne", "Dimension reduction method (t-sne)", false},
    {"neighbors-method", "nm", "vptree", "Neighbors search method (brute, vptree; the search is exact here)", false},
    {"eigen-method", "em", "dense", "Eigendecomposition method (unused by t-SNE)", false},
    {"computation-strategy", "cs", "cpu", "Computation strategy (unused by t-SNE)", false},
    {"target-dimension", "td", "2", "Target dimension", false},
    {"num-neighbors", "k", "10", "Number of neighbors (unused by t-SNE: it takes 3 x perplexity)", false},
    {"gaussian-width", "gw", "1.0", "Width of gaussian kernel (unused by t-SNE)", false},
    {"timesteps", "", "1", "Number of timesteps for diffusion map (unused by t-SNE)", false},
    {"max-iters", "", "1000", "Maximum number of iterations (t-SNE runs its fixed 1000)", false},
    {"sne-perplexity", "", "30.0", "Perplexity for the t-SNE algorithm", false},
    {"sne-theta", "", "0.5", "Theta for the t-SNE algorithm", false},
    {"seed", "", "", "Seed of std::rand for the initial embedding (default: time(NULL), as the reference)", false},
};

const std::vector<std::string> kTsneNames = {"t-sne", "t-stochastic_proximity_embedding"};
// include/tapkee/defines/method_names.hpp:15-57 (names of the other methods)
const std::vector<std::string> kOtherNames = {
    "local_tangent_space_alignment", "ltsa", "locally_linear_embedding", "lle", "hessian_locally_linear_embedding",
    "hlle", "multidimensional_scaling", "mds", "landmark_multidimensional_scaling", "l-mds", "isomap",
    "landmark_isomap", "l-isomap", "diffusion_map", "dm", "kernel_pca", "kpca", "pca", "random_projection", "ra",
    "laplacian_eigenmaps", "la", "locality_preserving_projections", "lpp", "neighborhood_preserving_embedding", "npe",
    "linear_local_tangent_space_alignment", "lltsa", "stochastic_proximity_embedding", "spe", "passthru",
    "factor_analysis", "fa", "manifold_sculpting"};

size_t edit_distance(const std::string& a, const std::string& b)
{
    std::vector<size_t> row(b.size() + 1);
    for (size_t j = 0; j <= b.size(); j++) row[j] = j;
    for (size_t i = 1; i <= a.size(); i++) {
        size_t diag = row[0];
        row[0] = i;
        for (size_t j = 1; j <= b.size(); j++) {
            const size_t up = row[j];
            row[j] = std::min({row[j] + 1, row[j - 1] + 1, diag + (a[i - 1] == b[j - 1] ? 0 : 1)});
            diag = up;
        }
    }
    return row[b.size()];
}

// Unknown value: say which known one is closest (util.hpp:171-189).
void suggest(const std::string& got, const std::vector<std::string>& known)
{
    auto best = std::min_element(known.begin(), known.end(), [&](const std::string& x, const std::string& y) {
        return edit_distance(got, x) < edit_distance(got, y);
    });
    if (best != known.end())
        tapkee::Logging::instance().message_info("Unknown parameter value `" + got + "`. Did you mean `" + *best + "`?");
}

void usage()
{
    std::cout << "tapkee (MI355X engine): t-distributed stochastic neighbor embedding\n\nOptions:\n";
    for (const Option& o : kOptions) {
        std::string names = (o.shrt.empty() ? "    " : "-" + o.shrt + ", ") + "--" + o.lng;
        std::cout << "  " << names << std::string(names.size() < 38 ? 38 - names.size() : 1, ' ') << o.help;
        if (!o.def.empty()) std::cout << " (default: " << o.def << ")";
        std::cout << "\n";
    }
}

// Lines of delimiter-separated values; every line must hold as many values
// as the first (util.hpp:73-122).
std::vector<std::vector<double>> read_rows(std::istream& in, char delimiter)
{
    std::vector<std::vector<double>> rows;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::vector<double> r;
        std::istringstream ls(line);
        std::string field;
        while (std::getline(ls, field, delimiter)) {
            std::istringstream fs(field);
            double v;
            if (fs >> v) r.push_back(v);
        }
        if (!rows.empty() && r.size() != rows[0].size())
            throw std::runtime_error("Wrong data at line " + std::to_string(rows.size()));
        rows.push_back(std::move(r));
    }
    return rows;
}

int run(int argc, const char** argv)
{
    std::map<std::string, std::string> val;
    std::map<std::string, bool> given;
    for (const Option& o : kOptions) val[o.lng] = o.def;
    for (int a = 1; a < argc; a++) {
        std::string arg = argv[a], name, value;
        bool has_value = false;
        if (arg.rfind("--", 0) == 0) {
            name = arg.substr(2);
        } else if (arg.size() > 1 && arg[0] == '-') {
            name = arg.substr(1);
        } else {
            throw std::runtime_error("unexpected argument " + arg);
        }
        const size_t eq = name.find('=');
        if (eq != std::string::npos) {
            value = name.substr(eq + 1);
            name = name.substr(0, eq);
            has_value = true;
        }
        const Option* opt = nullptr;
        for (const Option& o : kOptions)
            if (name == o.lng || (!o.shrt.empty() && name == o.shrt)) opt = &o;
        if (!opt) throw std::runtime_error("unknown option " + arg);
        if (!opt->flag && !has_value) {
            if (a + 1 >= argc) throw std::runtime_error("option " + arg + " needs a value");
            value = argv[++a];
        }
        val[opt->lng] = value;
        given[opt->lng] = true;
    }
    if (given["help"]) {
        usage();
        return 1;  // main.cpp:322-335 returns 1 after the usage text
    }
    tapkee::Logging& log = tapkee::Logging::instance();
    if (given["verbose"]) log.enable_info();
    if (given["debug"]) {
        log.enable_debug();
        log.message_info("Debug messages enabled");
    }
    if (given["benchmark"]) {
        log.enable_benchmark();
        log.message_info("Benchmarking enabled");
    }
    const std::string method = val["method"];
    if (std::find(kTsneNames.begin(), kTsneNames.end(), method) == kTsneNames.end()) {
        if (std::find(kOtherNames.begin(), kOtherNames.end(), method) != kOtherNames.end()) {
            log.message_error("Method " + method + " is not provided by this engine (only t-SNE runs on it)");
        } else {
            std::vector<std::string> all = kTsneNames;
            all.insert(all.end(), kOtherNames.begin(), kOtherNames.end());
            suggest(method, all);
            log.message_error("Unknown method " + method);
        }
        return 1;
    }
    const std::vector<std::string> nmethods = {"brute", "vptree", "covertree"};
    if (std::find(nmethods.begin(), nmethods.end(), val["neighbors-method"]) == nmethods.end()) {
        suggest(val["neighbors-method"], nmethods);
        log.message_error("Unknown neighbors method " + val["neighbors-method"]);
        return 1;
    }
    const int target_dim = std::stoi(val["target-dimension"]);
    if (target_dim <= 0) {
        log.message_error("\"Only\" a positive target dimensionality larger than zero is possible in current "
                          "circumstances. Please visit other universe");
        return 1;
    }
    if (std::stoi(val["num-neighbors"]) < 3) {
        log.message_error("The provided number of neighbors is too small, consider at least 3.");
        return 1;
    }
    if (std::stod(val["gaussian-width"]) < 0.0) {
        log.message_error("Width of the gaussian kernel is negative.");
        return 1;
    }
    if (std::stoi(val["timesteps"]) < 0) {
        log.message_error("Number of timesteps is negative.");
        return 1;
    }
    const char delimiter = val["delimiter"].empty() ? ',' : val["delimiter"][0];
    std::ifstream ifs(val["input-file"]);
    if (!ifs) throw std::runtime_error("cannot open " + val["input-file"]);
    const std::vector<std::vector<double>> rows = read_rows(ifs, delimiter);
    // features x samples, column-major (tapkee's DenseMatrix)
    const bool tin = given["transpose-input"];
    const int nr = (int)rows.size(), nc = rows.empty() ? 0 : (int)rows[0].size();
    const int N = tin ? nc : nr, D = tin ? nr : nc;
    tapkee::DenseMatrix X(D, N);
    for (int r = 0; r < nr; r++)
        for (int c = 0; c < nc; c++) {
            if (tin)
                X(r, c) = rows[r][c];
            else
                X(c, r) = rows[r][c];
        }
    log.message_info("Data contains " + std::to_string(N) + " feature vectors with dimension of " +
                     std::to_string(D));

    tapkee::EngineOptions eo;
    eo.opt.y0_mode = TKA_Y0_SEEDED;  // srand(seed), then the reference's draws
    eo.opt.seed = given["seed"] ? (uint32_t)std::stoul(val["seed"]) : (uint32_t)time(nullptr);
    eo.opt.verbose = 0;
    const tapkee::TapkeeOutput out =
        tapkee::with((tapkee::method = tapkee::tDistributedStochasticNeighborEmbedding,
                      tapkee::target_dimension = target_dim, tapkee::num_neighbors = std::stoi(val["num-neighbors"]),
                      tapkee::sne_perplexity = std::stod(val["sne-perplexity"]),
                      tapkee::sne_theta = std::stod(val["sne-theta"]), tapkee::engine_options = eo))
            .embedUsing(X);

    std::ofstream ofs(val["output-file"]);
    if (!ofs) throw std::runtime_error("cannot open " + val["output-file"]);
    const tapkee::DenseMatrix& E = out.embedding;  // N x target_dim
    const bool tout = given["transpose-output"];
    const int orow = tout ? E.cols() : E.rows(), ocol = tout ? E.rows() : E.cols();
    for (int i = 0; i < orow; i++) {
        for (int j = 0; j < ocol; j++) {
            ofs << (tout ? E(j, i) : E(i, j));
            if (j != ocol - 1) ofs << delimiter;
        }
        ofs << std::endl;
    }
    return 0;
}

}  // namespace

int main(int argc, const char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& exc) {
        std::cerr << "Some error occured: " << exc.what() << std::endl;  // main.cpp:550-563
        return 1;
    } catch (...) {
        std::cerr << "Unknown error occured" << std::endl;
        return 1;
    }
}
