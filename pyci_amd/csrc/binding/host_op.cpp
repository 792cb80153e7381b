// Sparse operator facade over the engine (SparseOp, sparseop.cpp).
#include <cstring>

#include "host.hpp"

namespace ph {

SparseOp::SparseOp(SQuantOp &ham, Wfn &wfn, i64 rows, i64 cols, bool symm, const std::string &storage_) {
    if (ham.nbasis != wfn.nbasis)
        throw py::value_error("hamiltonian and wave function have different nbasis");
    int st;
    if (storage_ == "csr") st = PYCI_STORAGE_CSR;
    else if (storage_ == "direct") st = PYCI_STORAGE_DIRECT;
    else if (storage_ == "auto") st = PYCI_STORAGE_AUTO;
    else throw py::value_error("storage must be 'auto', 'csr' or 'direct'");
    pyci_op *o = nullptr;
    check(pyci_op_create_ex(ham.device(), wfn.device(), rows, cols, symm ? 1 : 0, st, &o));
    op = std::shared_ptr<pyci_op>(o, pyci_op_destroy);
    refresh();
}

void SparseOp::refresh() {
    int64_t nr, nc, sz, nnz;
    int sym;
    double ec;
    check(pyci_op_info(op.get(), &nr, &nc, &sz, &nnz, &sym, &ec));
    nrow = nr;
    ncol = nc;
    size = sz;
    symmetric = sym != 0;
    ecore = ec;
    shape = py::make_tuple(nrow, ncol);
    int st = PYCI_STORAGE_CSR;
    check(pyci_op_storage(op.get(), &st));
    storage = st == PYCI_STORAGE_DIRECT ? "direct" : "csr";
}

void SparseOp::update(SQuantOp &ham, Wfn &wfn) {
    check(pyci_op_update(op.get(), ham.device(), wfn.device()));
    refresh();
}

Array<double> SparseOp::matvec(Array<double> x) const {
    if ((i64)x.size() != ncol) throw py::value_error("x must have ncol elements");
    Array<double> y(nrow);
    {
        py::gil_scoped_release nogil;
        check(pyci_op_matvec(op.get(), 1, x.data(), y.mutable_data()));
    }
    return y;
}

Array<double> SparseOp::matvec_out(Array<double> x, Array<double> out) const {
    if ((i64)x.size() != ncol) throw py::value_error("x must have ncol elements");
    if ((i64)out.size() != nrow) throw py::value_error("out must have nrow elements");
    {
        py::gil_scoped_release nogil;
        check(pyci_op_matvec(op.get(), 1, x.data(), out.mutable_data()));
    }
    return out;
}

double SparseOp::get_element(i64 i, i64 j) const {
    double v = 0.0;
    check(pyci_op_get_element(op.get(), i, j, &v));
    return v;
}

py::tuple SparseOp::solve(i64 n, py::object c0, i64 ncv, i64 maxiter, double tol, const std::string &precision) const {
    int prec;
    if (precision == "f64" || precision == "fp64" || precision == "double")
        prec = PYCI_PREC_F64;
    else if (precision == "mixed")
        prec = PYCI_PREC_MIXED;
    else
        throw py::value_error("precision must be 'f64' or 'mixed'");
    if (n < 1) throw py::value_error("n must be >= 1");
    Array<double> evals(n);
    Array<double> evecs({n, nrow});
    Array<double> cv;
    const double *cp = nullptr;
    if (!c0.is_none()) {
        cv = c0.cast<Array<double>>();
        if ((i64)cv.size() != nrow) throw py::value_error("c0 must have nrow elements");
        cp = cv.data();
    }
    pyci_solve_stats st;
    {
        py::gil_scoped_release nogil;
        check(pyci_op_solve(op.get(), n, cp, ncv, maxiter, tol, prec, evals.mutable_data(), evecs.mutable_data(),
                            &st));
    }
    return py::make_tuple(evals, evecs);
}

py::tuple SparseOp::export_csr() const {
    Array<double> data(size);
    Array<int64_t> indices(size);
    Array<int64_t> indptr(nrow + 1);
    check(pyci_op_export(op.get(), data.mutable_data(), indices.mutable_data(), indptr.mutable_data()));
    return py::make_tuple(data, indices, indptr);
}

Array<double> SparseOp::diagonal() const {
    Array<double> d(nrow);
    check(pyci_op_diagonal(op.get(), d.mutable_data()));
    return d;
}

}  // namespace ph
