// Command-line driver with the reference's flags and flow (src/main.cpp:13-96,
// parameters as src/Utilities.cpp:66-370 reads them): load the data and query
// text files, build the 1-D index, then query with qProbes = 1000, 2000, ...,
// 10000, each call verbose (statistics printed, ids written to
// 1D_Top_*.txt in the working directory).
//
//   falconnpp --n_points N --n_features D --X data.txt --n_queries Q --Q queries.txt
//             [--n_tables 10] [--n_proj max(256, 2^ceil(log2 D))] [--iProbes 1]
//             [--bucket_minSize 50] [--bucket_scale 1.0] [--n_threads -1]
//             [--random_seed -1] [--n_neighbors 1] [--qProbes 10] [--verbose]
//
// Addition of this driver: --qProbes_list "p1,p2,..." replaces the sweep with
// the given counts.  Errors end the program with status 1 and a message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/FalconnPP.hpp"

namespace {

// value of flag `name`, or nullptr
const char* arg(int nargs, char** args, const char* name) {
    for (int i = 1; i + 1 < nargs; ++i)
        if (std::strcmp(args[i], name) == 0) return args[i + 1];
    return nullptr;
}
bool has(int nargs, char** args, const char* name) {
    for (int i = 1; i < nargs; ++i)
        if (std::strcmp(args[i], name) == 0) return true;
    return false;
}

struct IndexParam {
    int n_points = 0, n_features = 0, bucket_minSize = 50, n_tables = 10, n_proj = 0, iProbes = 1, n_threads = -1,
        seed = -1;
    float bucket_scale = 1.0f;
};
struct QueryParam {
    int n_queries = 0, n_neighbors = 1, qProbes = 10;
    bool verbose = false;
};

[[noreturn]] void fail(const std::string& msg) {
    std::fprintf(stderr, "Error: %s\n", msg.c_str());
    std::exit(1);
}

// src/Utilities.cpp:66-255
IndexParam read_index_param(int nargs, char** args) {
    if (nargs < 6) fail("too few arguments");
    IndexParam p;
    const char* v;
    if (!(v = arg(nargs, args, "--n_points"))) fail("Number of rows/points is missing !");
    p.n_points = std::atoi(v);
    std::printf("Number of rows/points of X: %d\n", p.n_points);
    if (!(v = arg(nargs, args, "--n_features"))) fail("Number of columns/dimensions is missing !");
    p.n_features = std::atoi(v);
    std::printf("Number of columns/dimensions: %d\n", p.n_features);
    if ((v = arg(nargs, args, "--bucket_minSize"))) {
        p.bucket_minSize = std::atoi(v);
        std::printf("Minimum number of points in a bucket: %d\n", p.bucket_minSize);
    } else {
        std::printf("Default minimum number of points in a bucket: %d\n", p.bucket_minSize);
    }
    if ((v = arg(nargs, args, "--n_tables"))) {
        p.n_tables = std::atoi(v);
        std::printf("Number of LSH tables: %d\n", p.n_tables);
    } else {
        std::printf("Default number of LSH tables: %d\n", p.n_tables);
    }
    if ((v = arg(nargs, args, "--n_proj"))) {
        p.n_proj = std::atoi(v);
    } else {
        const int e = (int)std::ceil(std::log2(1.0 * p.n_features));
        p.n_proj = std::max(256, 1 << e);
    }
    std::printf("Number of projections: %d\n", p.n_proj);
    if ((v = arg(nargs, args, "--iProbes"))) {
        p.iProbes = std::atoi(v);
        std::printf("Number of index probes: %d\n", p.iProbes);
    } else {
        std::printf("Default index probes: %d\n", p.iProbes);
    }
    if ((v = arg(nargs, args, "--bucket_scale"))) {
        p.bucket_scale = (float)std::atof(v);
        std::printf("Bucket size scale: %g\n", p.bucket_scale);
    } else {
        std::printf("Default bucket scale: %g\n", p.bucket_scale);
    }
    if ((v = arg(nargs, args, "--n_threads"))) {
        p.n_threads = std::atoi(v);
        std::printf("Number of threads: %d\n", p.n_threads);
    }
    if ((v = arg(nargs, args, "--random_seed"))) {
        p.seed = std::atoi(v);
        std::printf("Random seed: %d\n", p.seed);
    } else {
        std::printf("Use a random seed !\n");
    }
    return p;
}

// src/Utilities.cpp:262-370
QueryParam read_query_param(int nargs, char** args) {
    if (nargs < 4) fail("too few arguments");
    QueryParam p;
    const char* v;
    if (!(v = arg(nargs, args, "--n_queries"))) fail("Number of queries is missing !");
    p.n_queries = std::atoi(v);
    std::printf("Number of rows/points of Q: %d\n", p.n_queries);
    if ((v = arg(nargs, args, "--n_neighbors"))) {
        p.n_neighbors = std::atoi(v);
        std::printf("Top-K query: %d\n", p.n_neighbors);
    } else {
        std::printf("Default top-K: %d\n", p.n_neighbors);
    }
    if ((v = arg(nargs, args, "--qProbes"))) {
        p.qProbes = std::atoi(v);
        std::printf("Number of query probes: %d\n", p.qProbes);
    } else {
        std::printf("Default query probes: %d\n", p.qProbes);
    }
    p.verbose = has(nargs, args, "--verbose");
    if (p.verbose) std::printf("Verbose mode: 1\n");
    return p;
}

std::vector<float> load(const char* flag, int nargs, char** args, int n, int d) {
    const char* path = arg(nargs, args, flag);
    if (!path || !*path) fail("Data file does not exist !");
    std::vector<float> m((size_t)n * d);
    if (fpp_load_txt(path, n, d, m.data(), nullptr) != FPP_OK) fail(fpp_last_error());
    std::printf("Finish reading data\n");
    return m;
}

}  // namespace

int main(int nargs, char** args) {
    const IndexParam ip = read_index_param(nargs, args);
    const QueryParam qp = read_query_param(nargs, args);
    if (ip.n_points < 1 || ip.n_features < 1 || qp.n_queries < 0) fail("n_points, n_features must be >= 1");
    const std::vector<float> X = load("--X", nargs, args, ip.n_points, ip.n_features);
    const std::vector<float> Q = load("--Q", nargs, args, qp.n_queries, ip.n_features);
    // the sweep of src/main.cpp:80-84, or the given list
    std::vector<int> probes;
    if (const char* list = arg(nargs, args, "--qProbes_list")) {
        for (const char* s = list; *s;) {
            char* e = nullptr;
            const long v = std::strtol(s, &e, 10);
            if (e == s) fail("bad --qProbes_list");
            probes.push_back((int)v);
            s = *e == ',' ? e + 1 : e;
        }
    } else {
        for (int i = 1; i <= 10; ++i) probes.push_back(1000 * i);
    }
    try {
        falconnpp::FalconnPP f(ip.n_points, ip.n_features);
        f.Index2Layers(ip.n_tables, ip.n_proj, ip.bucket_minSize, ip.bucket_scale, ip.iProbes, ip.n_threads, ip.seed);
        f.build2Layers_1D(X.data(), ip.n_features);
        std::printf("Finish building index... \n");
        for (int p : probes) {
            f.set_qProbes(p);
            f.query2Layers_1D(Q.data(), ip.n_features, qp.n_queries, qp.n_neighbors, qp.verbose);
        }
    } catch (const falconnpp::Error& e) {
        fail(e.what());
    }
    return 0;
}
