// pybind11 module "FalconnPP": the reference's Python surface
// (python/python_wrapper.cpp:14-30) over the MI355X engine.
//
// Same class name, method names, argument names and defaults.  Inputs are
// D x N / D x nq float32 matrices; like the reference's Eigen::Ref binding, an
// F-contiguous float32 array is used in place and anything else is converted.
// query returns an (nq, k) int32 array in Fortran order, as the reference's
// MatrixXi return does.  Errors raise ValueError (bad arguments or call order)
// or RuntimeError (device failures).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include "../../include/FalconnPP.hpp"

namespace py = pybind11;
using falconnpp::FalconnPP;

namespace {

using FArray = py::array_t<float, py::array::f_style | py::array::forcecast>;

void rethrow(const falconnpp::Error& e) {
    if (e.code == FPP_EINVAL || e.code == FPP_ESTATE) throw py::value_error(e.what());
    if (e.code == FPP_EIO) {
        PyErr_SetString(PyExc_OSError, e.what());
        throw py::error_already_set();
    }
    throw std::runtime_error(e.what());
}

FArray as_matrix(const py::handle& obj, const char* what) {
    FArray a = FArray::ensure(obj);
    if (!a) throw py::value_error(std::string(what) + " must be convertible to a float32 matrix");
    if (a.ndim() != 2) throw py::value_error(std::string(what) + " must be 2-D (n_features x count)");
    return a;
}

}  // namespace

PYBIND11_MODULE(FalconnPP, m) {
    m.doc() = "Falconn++ (cross-polytope LSF index) on AMD MI355X";
    py::class_<FalconnPP>(m, "FalconnPP")
        .def(py::init([](int n, int d) {
                 try {
                     return new FalconnPP(n, d);
                 } catch (const falconnpp::Error& e) {
                     rethrow(e);
                     return static_cast<FalconnPP*>(nullptr);
                 }
             }),
             py::arg("n_points"), py::arg("n_features"))
        .def(
            "setIndexParam",
            [](FalconnPP& f, int L, int D, int bucketLimit, float alpha, int p, int t, int s) {
                try {
                    f.Index2Layers(L, D, bucketLimit, alpha, p, t, s);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("n_tables"), py::arg("n_proj"), py::arg("bucket_minSize"), py::arg("bucket_scale"),
            py::arg("iProbes"), py::arg("n_threads") = 8, py::arg("random_seed") = -1)
        .def(
            "set_qProbes",
            [](FalconnPP& f, int p) {
                try {
                    f.set_qProbes(p);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("qProbes"))
        .def(
            "set_threads",
            [](FalconnPP& f, int t) {
                try {
                    f.set_threads(t);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("n_threads"))
        .def("clear",
             [](FalconnPP& f) {
                 try {
                     f.clear();
                 } catch (const falconnpp::Error& e) {
                     rethrow(e);
                 }
             })
        // streaming insert (no reference counterpart)
        .def(
            "set_insertable",
            [](FalconnPP& f, bool on) {
                try {
                    f.setInsertable(on);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("on") = true)
        .def(
            "insert",
            [](FalconnPP& f, const py::handle& pts) {
                FArray a = as_matrix(pts, "points");
                if (a.shape(0) != f.n_features()) throw py::value_error("points must be n_features x m");
                try {
                    py::gil_scoped_release nogil;
                    f.insert(a.data(), a.shape(0), (int)a.shape(1));
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("points"))
        // index files (no reference counterpart)
        .def(
            "save",
            [](FalconnPP& f, const std::string& path) {
                try {
                    py::gil_scoped_release nogil;
                    f.save(path);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("path"))
        .def(
            "load",
            [](FalconnPP& f, const std::string& path) {
                try {
                    py::gil_scoped_release nogil;
                    f.load(path);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("path"))
        .def(
            "build",
            [](FalconnPP& f, const py::handle& ds) {
                FArray a = as_matrix(ds, "dataset");
                if (a.shape(0) != f.n_features() || a.shape(1) < f.n_points())
                    throw py::value_error("dataset must be n_features x n_points (D x N)");
                try {
                    py::gil_scoped_release nogil;
                    f.build2Layers_1D(a.data(), a.shape(0));
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("dataset"))
        .def(
            "build2D",
            [](FalconnPP& f, const py::handle& ds) {
                FArray a = as_matrix(ds, "dataset");
                if (a.shape(0) != f.n_features() || a.shape(1) < f.n_points())
                    throw py::value_error("dataset must be n_features x n_points (D x N)");
                try {
                    py::gil_scoped_release nogil;
                    f.build2Layers(a.data(), a.shape(0));
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
            },
            py::arg("dataset"))
        .def(
            "query",
            [](FalconnPP& f, const py::handle& qs, int k, bool verbose) {
                FArray a = as_matrix(qs, "queries");
                if (a.shape(0) != f.n_features()) throw py::value_error("queries must be n_features x n_queries");
                const int nq = (int)a.shape(1);
                std::vector<int32_t> r;
                try {
                    py::gil_scoped_release nogil;
                    r = f.query2Layers_1D(a.data(), a.shape(0), nq, k, verbose);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
                py::array_t<int32_t, py::array::f_style> out({(py::ssize_t)nq, (py::ssize_t)k});
                auto w = out.mutable_unchecked<2>();
                for (int q = 0; q < nq; ++q)
                    for (int j = 0; j < k; ++j) w(q, j) = r[(size_t)q * k + j];
                return out;
            },
            py::arg("queries"), py::arg("n_neighbors"), py::arg("verbose") = false)
        .def(
            "query2D",
            [](FalconnPP& f, const py::handle& qs, int k, bool verbose) {
                FArray a = as_matrix(qs, "queries");
                if (a.shape(0) != f.n_features()) throw py::value_error("queries must be n_features x n_queries");
                const int nq = (int)a.shape(1);
                std::vector<int32_t> r;
                try {
                    py::gil_scoped_release nogil;
                    r = f.query2Layers(a.data(), a.shape(0), nq, k, verbose);
                } catch (const falconnpp::Error& e) {
                    rethrow(e);
                }
                py::array_t<int32_t, py::array::f_style> out({(py::ssize_t)nq, (py::ssize_t)k});
                auto w = out.mutable_unchecked<2>();
                for (int q = 0; q < nq; ++q)
                    for (int j = 0; j < k; ++j) w(q, j) = r[(size_t)q * k + j];
                return out;
            },
            py::arg("queries"), py::arg("n_neighbors"), py::arg("verbose") = false);
}
