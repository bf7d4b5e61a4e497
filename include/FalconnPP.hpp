// C++ host class of the MI355X Falconn++ engine, with the member names of the
// reference class FalconnPP (src/FalconnPP.h:10-132).  It is a thin owner of
// an fpp_index from the C ABI (falconnpp_amd.h); every failure becomes a
// falconnpp::Error carrying the C status code.
//
// Matrices are passed as (pointer, leading dimension) in the reference's
// column-major D x N layout instead of Eigen::Ref<const MatrixXf>.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "falconnpp_amd.h"

namespace falconnpp {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != FPP_OK) throw Error(rc, fpp_last_error());
}

class FalconnPP {
public:
    // src/FalconnPP.h:80-83
    FalconnPP(int n, int d) : n_points_(n), n_features_(d) { check(fpp_create(n, d, &h_)); }
    ~FalconnPP() { fpp_destroy(h_); }
    FalconnPP(const FalconnPP&) = delete;
    FalconnPP& operator=(const FalconnPP&) = delete;

    // src/FalconnPP.h:85-100
    void Index2Layers(int L, int D, int bucketLimit, float alpha, int p, int t, int s) {
        check(fpp_set_index_param(h_, L, D, bucketLimit, alpha, p, t, s));
    }
    // src/FalconnPP.h:102-123
    void clear() { check(fpp_clear(h_)); }
    void set_qProbes(int p) { check(fpp_set_qprobes(h_, p)); }
    void set_threads(int t) { check(fpp_set_threads(h_, t)); }

    // src/FalconnPP.cpp:21-303: X is D x N column-major, leading dimension ld
    void build2Layers_1D(const float* X, int64_t ld) { check(fpp_build(h_, X, ld)); }
    // src/FalconnPP.cpp:310-707: returns nq x k ids, row-major (query-major)
    std::vector<int32_t> query2Layers_1D(const float* Q, int64_t ldq, int nq, int k, bool verbose = false) {
        std::vector<int32_t> out((size_t)nq * k, 0);
        check(fpp_query(h_, Q, ldq, nq, k, out.data(), verbose ? 1 : 0));
        return out;
    }
    // The reference's 2D-storage variants (src/FalconnPP.cpp:722-1413) keep the
    // same bucket contents and return the same ids; they are served from the
    // same CSR tables.  A verbose query2Layers writes 2D_Top_...txt.
    void build2Layers(const float* X, int64_t ld) { build2Layers_1D(X, ld); }
    std::vector<int32_t> query2Layers(const float* Q, int64_t ldq, int nq, int k, bool verbose = false) {
        std::vector<int32_t> out((size_t)nq * k, 0);
        check(fpp_query_2d(h_, Q, ldq, nq, k, out.data(), verbose ? 1 : 0));
        return out;
    }

    // streaming insert (no reference counterpart; src/FalconnPP.cpp:711-713
    // announces it): setInsertable(true) before build, then insert appends m
    // points (D x m column-major, leading dimension ld) as ids n .. n+m-1
    void setInsertable(bool on) { check(fpp_set_insertable(h_, on ? 1 : 0)); }
    void insert(const float* X, int64_t ld, int m) {
        check(fpp_insert(h_, X, ld, m));
        n_points_ += m;
    }

    // index files (no reference counterpart)
    void save(const std::string& path) { check(fpp_save(h_, path.c_str())); }
    void load(const std::string& path) { check(fpp_load(h_, path.c_str())); }

    int n_points() const { return n_points_; }
    int n_features() const { return n_features_; }
    fpp_index* handle() const { return h_; }

private:
    fpp_index* h_ = nullptr;
    int n_points_, n_features_;
};

}  // namespace falconnpp
