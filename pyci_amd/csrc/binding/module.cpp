// pybind11 module `pyci_amd._pyci`: the reference's `pyci._pyci` surface
// (pyci/src/binding.cpp:29-1507) for the integrals, wave functions, sparse
// operator, heat-bath selection and RDMs, backed by the GPU engine.
#include <cstring>

#include "host.hpp"

namespace ph {
static i64 g_threads = 1;

template <class W>
static void copy_from(W &dst, const Wfn &src) {
    std::vector<u64> buf = src.hdets();
    dst.set_dets(std::move(buf), src.ndet);
    dst.full_colex = src.full_colex;
}

template <class W>
static std::unique_ptr<W> from_dets(i64 nb, i64 nu, i64 nd, Array<u64> a) {
    auto w = std::make_unique<W>(nb, nu, nd);
    i64 n = a.ndim() ? a.shape(0) : 0;
    if ((i64)a.size() != n * w->nwdet) throw py::value_error("determinant array has the wrong shape");
    w->set_dets(std::vector<u64>(a.data(), a.data() + n * w->nwdet), n);
    return w;
}

template <class W>
static std::unique_ptr<W> from_occs(i64 nb, i64 nu, i64 nd, Array<i64> a) {
    auto w = std::make_unique<W>(nb, nu, nd);
    const bool two = (w->kind == PYCI_WFN_FULLCI);
    const i64 per = two ? 2 * nu : nu;
    i64 n = a.ndim() ? a.shape(0) : 0;
    if ((i64)a.size() != n * per) throw py::value_error("occupation array has the wrong shape");
    std::vector<u64> buf((size_t)(n * w->nwdet), 0ULL);
    const i64 *p = a.data();
    for (i64 i = 0; i < n; ++i) {
        u64 *d = &buf[(size_t)(i * w->nwdet)];
        fill_det(nu, p + i * per, d);
        if (two) fill_det(nd, p + i * per + nu, d + w->nword);
    }
    w->set_dets(std::move(buf), n);
    return w;
}

template <class W>
static std::unique_ptr<W> from_file(const std::string &fn) {
    auto w = std::make_unique<W>(1, 0, 0);
    w->load_file(fn);
    return w;
}

// FullCI from DOCI (fullciwfn.cpp:22-33): alpha string = beta string.
static std::unique_ptr<FullCIWfn> fullci_from_doci(const DOCIWfn &d) {
    auto w = std::make_unique<FullCIWfn>(d.nbasis, d.nocc_up, d.nocc_dn);
    std::vector<u64> buf((size_t)(d.ndet * w->nwdet));
    for (i64 i = 0; i < d.ndet; ++i) {
        std::memcpy(&buf[(size_t)(i * w->nwdet)], d.det_ptr(i), sizeof(u64) * d.nword);
        std::memcpy(&buf[(size_t)(i * w->nwdet + d.nword)], d.det_ptr(i), sizeof(u64) * d.nword);
    }
    w->set_dets(std::move(buf), d.ndet);
    return w;
}

// GenCI from FullCI (genciwfn.cpp:26-40): spin-orbital p (alpha) and nbasis+q
// (beta). The reference fills the beta occupations from the alpha string
// (genciwfn.cpp:37); this conversion uses the beta string.
static std::unique_ptr<GenCIWfn> genci_from_fullci(const FullCIWfn &f) {
    auto w = std::make_unique<GenCIWfn>(f.nbasis * 2, f.nocc, 0);
    std::vector<u64> buf((size_t)(f.ndet * w->nwdet), 0ULL);
    std::vector<i64> occ((size_t)f.nocc + 1);
    for (i64 i = 0; i < f.ndet; ++i) {
        fill_occs(f.nword, f.det_ptr(i), occ.data());
        fill_occs(f.nword, f.det_ptr(i) + f.nword, occ.data() + f.nocc_up);
        for (i64 j = 0; j < f.nocc_dn; ++j) occ[(size_t)(f.nocc_up + j)] += f.nbasis;
        fill_det(f.nocc, occ.data(), &buf[(size_t)(i * w->nwdet)]);
    }
    w->set_dets(std::move(buf), f.ndet);
    return w;
}

static py::tuple det_array_range(const Wfn &w, i64 &lo, i64 &hi) {
    if (lo == -1) {
        lo = 0;
        if (hi == -1) hi = w.ndet;
    } else if (hi == -1) {
        hi = lo;
        lo = 0;
    }
    if (lo < 0 || hi > w.ndet || lo > hi) throw py::index_error("determinant range out of bounds");
    return py::tuple();
}

static Array<u64> to_det_array(const Wfn &w, i64 lo, i64 hi) {
    det_array_range(w, lo, hi);
    const bool two = (w.kind == PYCI_WFN_FULLCI);
    Array<u64> a = two ? Array<u64>({hi - lo, (i64)2, w.nword}) : Array<u64>({hi - lo, w.nword});
    if (hi > lo) std::memcpy(a.mutable_data(), w.det_ptr(lo), sizeof(u64) * (size_t)((hi - lo) * w.nwdet));
    return a;
}

static Array<i64> to_occ_array(const Wfn &w, i64 lo, i64 hi) {
    det_array_range(w, lo, hi);
    const bool two = (w.kind == PYCI_WFN_FULLCI);
    Array<i64> a = two ? Array<i64>({hi - lo, (i64)2, w.nocc_up}) : Array<i64>({hi - lo, w.nocc_up});
    i64 *p = a.mutable_data();
    const i64 per = two ? 2 * w.nocc_up : w.nocc_up;
    std::fill(p, p + (hi - lo) * per, 0);
    for (i64 i = lo; i < hi; ++i) {
        fill_occs(w.nword, w.det_ptr(i), p + (i - lo) * per);
        if (two) fill_occs(w.nword, w.det_ptr(i) + w.nword, p + (i - lo) * per + w.nocc_up);
    }
    return a;
}

static Array<u64> getitem(const Wfn &w, i64 i) {
    if (i < 0) i += w.ndet;
    if (i < 0 || i >= w.ndet) throw py::index_error("determinant index out of range");
    const bool two = (w.kind == PYCI_WFN_FULLCI);
    Array<u64> a = two ? Array<u64>({(i64)2, w.nword}) : Array<u64>(w.nword);
    std::memcpy(a.mutable_data(), w.det_ptr(i), sizeof(u64) * (size_t)w.nwdet);
    return a;
}

static const u64 *det_arg(const Wfn &w, const Array<u64> &d) {
    if ((i64)d.size() != w.nwdet) throw py::value_error("determinant has the wrong number of words");
    return d.data();
}

static void add_dets_from(Wfn &dst, const Wfn &src) {
    if (src.nwdet != dst.nwdet) throw py::value_error("incompatible wave functions");
    for (i64 i = 0; i < src.ndet; ++i) dst.add_det(src.det_ptr(i));
}

// Learned selector (no counterpart in the reference): the network of
// pyci_mlp_* over the determinants' occupation bits.
struct MLPSelector {
    std::shared_ptr<pyci_mlp> net;
    i64 nfeat = 0;
    MLPSelector(i64 nf, u64 seed) : nfeat(nf) {
        pyci_mlp *m = nullptr;
        check(pyci_mlp_create(context(), nf, seed, &m));
        net.reset(m, [](pyci_mlp *p) { pyci_mlp_destroy(p); });
    }
    Array<float> get_params() const {
        Array<float> p((i64)PYCI_MLP_NPARAM);
        check(pyci_mlp_get_params(net.get(), p.mutable_data()));
        return p;
    }
    void set_params(Array<float> p) {
        if (p.size() != PYCI_MLP_NPARAM) throw py::value_error("expected PYCI_MLP_NPARAM parameters");
        check(pyci_mlp_set_params(net.get(), p.data()));
    }
    Array<float> predict(const Wfn &w) const {
        Array<float> y(w.ndet);
        if (w.ndet == 0) return y;
        py::gil_scoped_release nogil;
        check(pyci_mlp_predict(net.get(), w.kind, w.nbasis, w.ndet, w.hdets().data(), y.mutable_data()));
        return y;
    }
    py::tuple loss_grad(const Wfn &w, Array<float> target) const {
        if ((i64)target.size() != w.ndet) throw py::value_error("target must have one entry per determinant");
        Array<float> g((i64)PYCI_MLP_NPARAM);
        float loss = 0.0f;
        {
            py::gil_scoped_release nogil;
            check(pyci_mlp_loss_grad(net.get(), w.kind, w.nbasis, w.ndet, w.hdets().data(), target.data(), &loss,
                                     g.mutable_data()));
        }
        return py::make_tuple(loss, g);
    }
    float train(Wfn &w, Array<double> coeffs, i64 steps, i64 batch, double lr, double cmin, u64 seed) {
        if ((i64)coeffs.size() != w.ndet) throw py::value_error("coeffs must have one entry per determinant");
        pyci_wfn *dw = w.device();
        float loss = 0.0f;
        py::gil_scoped_release nogil;
        check(pyci_mlp_train(net.get(), dw, coeffs.data(), cmin, steps, batch, lr, seed, &loss));
        return loss;
    }
};

template <class W>
static i64 add_hci(SQuantOp &ham, W &wfn, Array<double> coeffs, double eps, i64, py::object selector, double pool,
                   i64 nadd, int score_mode) {
    if ((i64)coeffs.size() != wfn.ndet) throw py::value_error("coeffs must have one entry per determinant");
    if (ham.nbasis != wfn.nbasis) throw py::value_error("hamiltonian and wave function have different nbasis");
    MLPSelector *sel = selector.is_none() ? nullptr : selector.cast<MLPSelector *>();
    pyci_wfn *dw = wfn.device();
    pyci_ham *dh = ham.device();
    const i64 old = wfn.ndet;
    int64_t added = 0;
    {
        py::gil_scoped_release nogil;
        if (sel)
            check(pyci_add_hci_mlp(dh, dw, coeffs.data(), eps, sel->net.get(), pool, nadd, score_mode, &added));
        else
            check(pyci_add_hci(dh, dw, coeffs.data(), eps, &added));
    }
    wfn.adopt_device_appends(old);
    return added;
}

template <class W>
static double enpt2(SQuantOp &ham, W &wfn, Array<double> coeffs, double energy, double eps, i64) {
    if ((i64)coeffs.size() != wfn.ndet) throw py::value_error("coeffs must have one entry per determinant");
    if (ham.nbasis != wfn.nbasis) throw py::value_error("hamiltonian and wave function have different nbasis");
    pyci_wfn *dw = wfn.device();
    pyci_ham *dh = ham.device();
    double e = 0.0;
    {
        py::gil_scoped_release nogil;
        check(pyci_compute_enpt2(dh, dw, coeffs.data(), energy, eps, &e, nullptr));
    }
    return e;
}

static py::tuple rdms(Wfn &wfn, Array<double> coeffs) {
    if ((i64)coeffs.size() != wfn.ndet) throw py::value_error("coeffs must have one entry per determinant");
    const i64 n = wfn.nbasis;
    Array<double> r1, r2;
    if (wfn.kind == PYCI_WFN_DOCI) {
        r1 = Array<double>({n, n});
        r2 = Array<double>({n, n});
    } else if (wfn.kind == PYCI_WFN_FULLCI) {
        r1 = Array<double>({(i64)2, n, n});
        r2 = Array<double>({(i64)3, n, n, n, n});
    } else {
        r1 = Array<double>({n, n});
        r2 = Array<double>({n, n, n, n});
    }
    pyci_wfn *dw = wfn.device();
    {
        py::gil_scoped_release nogil;
        check(pyci_compute_rdms(dw, coeffs.data(), r1.mutable_data(), r2.mutable_data()));
    }
    return py::make_tuple(r1, r2);
}

static py::tuple trdms(Wfn &w1, Wfn &w2, Array<double> c1, Array<double> c2) {
    if ((i64)c1.size() != w1.ndet || (i64)c2.size() != w2.ndet)
        throw py::value_error("coeffs must have one entry per determinant");
    if (w1.kind != w2.kind || w1.nbasis != w2.nbasis || w1.nocc_up != w2.nocc_up || w1.nocc_dn != w2.nocc_dn)
        throw py::value_error("wave functions must have the same kind, nbasis and occupations");
    const i64 n = w1.nbasis;
    Array<double> r1, r2;
    if (w1.kind == PYCI_WFN_DOCI) {
        r1 = Array<double>({n, n});
        r2 = Array<double>({n, n});
    } else if (w1.kind == PYCI_WFN_FULLCI) {
        r1 = Array<double>({(i64)2, n, n});
        r2 = Array<double>({(i64)3, n, n, n, n});
    } else {
        r1 = Array<double>({n, n});
        r2 = Array<double>({n, n, n, n});
    }
    pyci_wfn *d1 = w1.device();
    pyci_wfn *d2 = w2.device();
    {
        py::gil_scoped_release nogil;
        check(pyci_compute_transition_rdms(d1, d2, c1.data(), c2.data(), r1.mutable_data(), r2.mutable_data()));
    }
    return py::make_tuple(r1, r2);
}

// FanCI geminal objectives (AP1roGObjective / APIGObjective, fanci.cpp,
// ap1rog.cpp, apig.cpp): overlaps, objective and Jacobian on the GPU.
struct FanciObjective {
    // declared before f so that it is destroyed after it: the objective's handle
    // refers to the operator until pyci_fanci_destroy returns
    std::shared_ptr<pyci_op> op;  // the operator the objective was built over (kept alive)
    std::shared_ptr<pyci_fanci> f;
    i64 nparam = 0, nproj = 0, nconn = 0, nequation = 0;

    FanciObjective(int kind, SparseOp &o, DOCIWfn &w, py::object idc, py::object vdc, py::object ipc,
                   py::object vpc) {
        if (idc.is_none() != vdc.is_none() || ipc.is_none() != vpc.is_none())
            throw py::value_error("constraint indices and values must be given together");
        Array<i64> a_idc = idc.is_none() ? Array<i64>(0) : idc.cast<Array<i64>>();
        Array<double> a_vdc = vdc.is_none() ? Array<double>(0) : vdc.cast<Array<double>>();
        Array<i64> a_ipc = ipc.is_none() ? Array<i64>(0) : ipc.cast<Array<i64>>();
        Array<double> a_vpc = vpc.is_none() ? Array<double>(0) : vpc.cast<Array<double>>();
        if (a_idc.size() != a_vdc.size() || a_ipc.size() != a_vpc.size())
            throw py::value_error("constraint indices and values differ in length");
        pyci_fanci *h = nullptr;
        check(pyci_fanci_create(o.op.get(), w.device(), kind, (i64)a_idc.size(), a_idc.size() ? a_idc.data() : nullptr,
                                a_vdc.size() ? a_vdc.data() : nullptr, (i64)a_ipc.size(),
                                a_ipc.size() ? a_ipc.data() : nullptr, a_vpc.size() ? a_vpc.data() : nullptr, &h));
        f = std::shared_ptr<pyci_fanci>(h, pyci_fanci_destroy);
        op = o.op;
        int64_t np, nr, nc, ne;
        check(pyci_fanci_info(h, &np, &nr, &nc, &ne));
        nparam = np, nproj = nr, nconn = nc, nequation = ne;
    }
    void same_op(const SparseOp &o) const {
        if (o.op.get() != op.get()) throw py::value_error("the objective was built over another operator");
    }
    void check_x(const Array<double> &x, i64 n) const {
        if ((i64)x.size() < n) throw py::value_error("parameter array is too short");
    }
    Array<double> overlap(Array<double> x) const {
        check_x(x, nparam);
        Array<double> y(nconn);
        py::gil_scoped_release nogil;
        check(pyci_fanci_overlap(f.get(), x.data(), y.mutable_data()));
        return y;
    }
    py::array_t<double, py::array::f_style> d_overlap(Array<double> x) const {
        check_x(x, nparam);
        py::array_t<double, py::array::f_style> y({nconn, nparam});
        py::gil_scoped_release nogil;
        check(pyci_fanci_d_overlap(f.get(), x.data(), y.mutable_data()));
        return y;
    }
    Array<double> objective(const SparseOp &o, Array<double> x) const {
        same_op(o);
        check_x(x, nparam + 1);
        Array<double> y(nequation);
        py::gil_scoped_release nogil;
        check(pyci_fanci_objective(f.get(), x.data(), y.mutable_data()));
        return y;
    }
    py::array_t<double, py::array::f_style> jacobian(const SparseOp &o, Array<double> x) const {
        same_op(o);
        check_x(x, nparam + 1);
        py::array_t<double, py::array::f_style> y({nequation, nparam + 1});
        py::gil_scoped_release nogil;
        check(pyci_fanci_jacobian(f.get(), x.data(), y.mutable_data()));
        return y;
    }
};

// one Python class per objective kind (pybind11 registers each C++ type once)
template <int KIND>
struct FanciKind : FanciObjective {
    FanciKind(SparseOp &o, DOCIWfn &w, py::object a, py::object b, py::object c, py::object d)
        : FanciObjective(KIND, o, w, a, b, c, d) {}
};

template <int KIND>
static void bind_fanci(py::module_ &m, const char *name) {
    using C = FanciKind<KIND>;
    py::class_<C>(m, name)
        .def(py::init<SparseOp &, DOCIWfn &, py::object, py::object, py::object, py::object>(), py::arg("op"),
             py::arg("wfn"), py::arg("idx_det_cons") = py::none(), py::arg("det_cons") = py::none(),
             py::arg("idx_param_cons") = py::none(), py::arg("param_cons") = py::none())
        .def_readonly("nparam", &C::nparam)
        .def_readonly("nproj", &C::nproj)
        .def_readonly("nconn", &C::nconn)
        .def_readonly("nequation", &C::nequation)
        .def("overlap", &C::overlap, py::arg("x"))
        .def("d_overlap", &C::d_overlap, py::arg("x"))
        .def("objective", &C::objective, py::arg("op"), py::arg("x"))
        .def("jacobian", &C::jacobian, py::arg("op"), py::arg("x"));
}

}  // namespace ph

using namespace ph;

PYBIND11_MODULE(_pyci, m) {
    py::options options;
    options.disable_function_signatures();
    m.doc() = "pyci_amd: MI355X (gfx950) configuration-interaction engine behind PyCI's _pyci API";
    m.attr("__version__") = "0.6.1+mi355x";
    m.attr("c_long") = py::dtype::of<int64_t>();
    m.attr("c_ulong") = py::dtype::of<uint64_t>();
    m.attr("c_double") = py::dtype::of<double>();

    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const EngineTypeError &e) {
            PyErr_SetString(PyExc_TypeError, e.what());
        }
    });

    // ---- integrals ----
    py::class_<SQuantOp> sq(m, "secondquant_op");
    sq.def_readonly("nbasis", &SQuantOp::nbasis)
        .def_readonly("ecore", &SQuantOp::ecore)
        .def_readonly("one_mo", &SQuantOp::one_mo)
        .def_readonly("two_mo", &SQuantOp::two_mo)
        .def_readonly("h", &SQuantOp::h)
        .def_readonly("v", &SQuantOp::v)
        .def_readonly("w", &SQuantOp::w)
        .def(py::init<const std::string &>(), py::arg("filename"))
        .def(py::init<double, Array<double>, Array<double>>(), py::arg("ecore"), py::arg("one_mo"), py::arg("two_mo"))
        .def("to_file", &SQuantOp::to_file, py::arg("filename"), py::arg("nelec") = 0, py::arg("ms2") = 0,
             py::arg("tol") = 0.0);

    // ---- wave functions ----
    py::class_<Wfn> wf(m, "wavefunction");
    wf.def_readonly("nbasis", &Wfn::nbasis)
        .def_readonly("nocc", &Wfn::nocc)
        .def_readonly("nocc_up", &Wfn::nocc_up)
        .def_readonly("nocc_dn", &Wfn::nocc_dn)
        .def_readonly("nvir", &Wfn::nvir)
        .def_readonly("nvir_up", &Wfn::nvir_up)
        .def_readonly("nvir_dn", &Wfn::nvir_dn)
        .def("__len__", &Wfn::length)
        .def("squeeze", &Wfn::squeeze)
        // determinants that add_hci appended on the device and host code has not
        // read yet (they stay on the device across update and solve)
        .def_property_readonly("device_only_dets", &Wfn::host_pending);

    auto common = [](auto &cls) {
        using C = typename std::remove_reference_t<decltype(cls)>::type;
        cls.def("__getitem__", [](const C &w, i64 i) { return getitem(w, i); }, py::arg("index"))
            .def("to_file", [](const C &w, const std::string &f) { w.to_file(f); }, py::arg("filename"))
            .def("to_det_array", [](const C &w, i64 lo, i64 hi) { return to_det_array(w, lo, hi); },
                 py::arg("low") = -1, py::arg("high") = -1)
            .def("to_occ_array", [](const C &w, i64 lo, i64 hi) { return to_occ_array(w, lo, hi); },
                 py::arg("low") = -1, py::arg("high") = -1)
            .def("index_det", [](const C &w, Array<u64> d) { return w.index_det(det_arg(w, d)); }, py::arg("det"))
            .def("index_det_from_rank",
                 [](const C &w, std::pair<u64, u64> r) { return w.index_det_from_rank(r); }, py::arg("rank"))
            .def("rank_det", [](const C &w, Array<u64> d) { return w.rank_det(det_arg(w, d)); }, py::arg("det"))
            .def("add_det", [](C &w, Array<u64> d) { return w.add_det(det_arg(w, d)); }, py::arg("det"))
            .def("add_occs",
                 [](C &w, Array<i64> o) {
                     const i64 need = (w.kind == PYCI_WFN_FULLCI) ? w.nocc_up + w.nocc_dn : w.nocc_up;
                     if ((i64)o.size() < need) throw py::value_error("occs has too few entries");
                     return w.add_det_from_occs(o.data());
                 },
                 py::arg("occs"))
            .def("add_hartreefock_det", &C::add_hartreefock_det)
            .def("add_all_dets", [](C &w, i64) { w.fill_all_dets(); }, py::arg("nthread") = -1)
            .def("add_excited_dets", &C::py_add_excited_dets, py::arg("exc"), py::arg("ref") = py::none())
            .def("add_dets_from_wfn", [](C &w, const C &o) { add_dets_from(w, o); }, py::arg("wfn"))
            .def("reserve", [](C &w, i64 n) { w.reserve(n); }, py::arg("n"));
    };
    py::class_<OneSpinWfn, Wfn> one(m, "one_spin_wfn");
    common(one);
    py::class_<TwoSpinWfn, Wfn> two(m, "two_spin_wfn");
    common(two);

    py::class_<DOCIWfn, OneSpinWfn> doci(m, "doci_wfn");
    doci.def(py::init([](const DOCIWfn &o) {
                 auto w = std::make_unique<DOCIWfn>(o.nbasis, o.nocc_up, o.nocc_dn);
                 copy_from(*w, o);
                 return w;
             }),
             py::arg("wfn"))
        .def(py::init([](const std::string &f) {
                 auto w = from_file<DOCIWfn>(f);
                 if (w->nocc_up != w->nocc_dn) throw std::invalid_argument("nocc_up != nocc_dn");
                 w->kind = PYCI_WFN_DOCI;
                 return w;
             }),
             py::arg("filename"))
        .def(py::init<i64, i64, i64>(), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"))
        .def(py::init(&from_dets<DOCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"))
        .def(py::init(&from_occs<DOCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"));

    py::class_<FullCIWfn, TwoSpinWfn> fci(m, "fullci_wfn");
    fci.def(py::init(&fullci_from_doci), py::arg("wfn"))
        .def(py::init([](const FullCIWfn &o) {
                 auto w = std::make_unique<FullCIWfn>(o.nbasis, o.nocc_up, o.nocc_dn);
                 copy_from(*w, o);
                 return w;
             }),
             py::arg("wfn"))
        .def(py::init(&from_file<FullCIWfn>), py::arg("filename"))
        .def(py::init<i64, i64, i64>(), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"))
        .def(py::init(&from_dets<FullCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"))
        .def(py::init(&from_occs<FullCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"));

    py::class_<GenCIWfn, OneSpinWfn> gen(m, "genci_wfn");
    gen.def(py::init([](const DOCIWfn &d) { return genci_from_fullci(*fullci_from_doci(d)); }), py::arg("wfn"))
        .def(py::init([](const FullCIWfn &f) { return genci_from_fullci(f); }), py::arg("wfn"))
        .def(py::init([](const GenCIWfn &o) {
                 auto w = std::make_unique<GenCIWfn>(o.nbasis, o.nocc_up, o.nocc_dn);
                 copy_from(*w, o);
                 return w;
             }),
             py::arg("wfn"))
        .def(py::init([](const std::string &f) {
                 auto w = from_file<GenCIWfn>(f);
                 if (w->nocc_dn) throw std::invalid_argument("nocc_dn != 0");
                 return w;
             }),
             py::arg("filename"))
        .def(py::init<i64, i64, i64>(), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"))
        .def(py::init(&from_dets<GenCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"))
        .def(py::init(&from_occs<GenCIWfn>), py::arg("nbasis"), py::arg("nocc_up"), py::arg("nocc_dn"),
             py::arg("array"));

    // ---- sparse operator ----
    py::class_<SparseOp> so(m, "sparse_op");
    so.def_readonly("ecore", &SparseOp::ecore)
        .def_readonly("symmetric", &SparseOp::symmetric)
        .def_readonly("size", &SparseOp::size)
        .def_readonly("shape", &SparseOp::shape)
        .def_readonly("storage", &SparseOp::storage)
        .def_property_readonly("dtype", [](const SparseOp &) { return py::dtype::of<double>(); });
    so.def(py::init([](SQuantOp &h, DOCIWfn &w, i64 r, i64 c, bool s, const std::string &st) {
                 return new SparseOp(h, w, r, c, s, st);
             }),
           py::arg("ham"), py::arg("wfn"), py::arg("nrow") = -1, py::arg("ncol") = -1, py::arg("symmetric") = true,
             py::arg("storage") = "auto")
        .def(py::init([](SQuantOp &h, FullCIWfn &w, i64 r, i64 c, bool s, const std::string &st) {
                 return new SparseOp(h, w, r, c, s, st);
             }),
             py::arg("ham"), py::arg("wfn"), py::arg("nrow") = -1, py::arg("ncol") = -1, py::arg("symmetric") = true,
             py::arg("storage") = "auto")
        .def(py::init([](SQuantOp &h, GenCIWfn &w, i64 r, i64 c, bool s, const std::string &st) {
                 return new SparseOp(h, w, r, c, s, st);
             }),
             py::arg("ham"), py::arg("wfn"), py::arg("nrow") = -1, py::arg("ncol") = -1, py::arg("symmetric") = true,
             py::arg("storage") = "auto")
        .def("update", [](SparseOp &o, SQuantOp &h, DOCIWfn &w) { o.update(h, w); }, py::arg("ham"), py::arg("wfn"))
        .def("update", [](SparseOp &o, SQuantOp &h, FullCIWfn &w) { o.update(h, w); }, py::arg("ham"), py::arg("wfn"))
        .def("update", [](SparseOp &o, SQuantOp &h, GenCIWfn &w) { o.update(h, w); }, py::arg("ham"), py::arg("wfn"))
        .def("__call__", &SparseOp::matvec, py::arg("x"))
        .def("__call__", &SparseOp::matvec_out, py::arg("x"), py::arg("out"))
        .def("matvec", &SparseOp::matvec, py::arg("x"))
        .def("matvec", &SparseOp::matvec_out, py::arg("x"), py::arg("out"))
        .def("get_element", &SparseOp::get_element, py::arg("i"), py::arg("j"))
        .def("solve", &SparseOp::solve, py::arg("n") = 1, py::arg("c0") = py::none(), py::arg("ncv") = -1,
             py::arg("maxiter") = -1, py::arg("tol") = 1.0e-12, py::arg("precision") = "f64")
        .def("reserve", [](SparseOp &, i64) {}, py::arg("n"))
        .def("squeeze", [](SparseOp &) {})
        .def("data", [](const SparseOp &o) { return py::object(o.export_csr()[0]); })
        .def("indices", [](const SparseOp &o) { return py::object(o.export_csr()[1]); })
        .def("indptr", [](const SparseOp &o) { return py::object(o.export_csr()[2]); })
        .def("export", &SparseOp::export_csr)
        .def("diagonal", &SparseOp::diagonal);

    // ---- FanCI geminal objectives ----
    bind_fanci<PYCI_FANCI_AP1ROG>(m, "AP1roGObjective");
    bind_fanci<PYCI_FANCI_APIG>(m, "APIGObjective");

    // ---- free functions ----
    m.def("get_num_threads", [] { return g_threads; });
    m.def("set_num_threads", [](i64 n) { g_threads = std::max<i64>(n, 1); }, py::arg("n"));
    m.def("popcnt",
          [](Array<u64> d) {
              i64 c = 0;
              for (i64 i = 0; i < (i64)d.size(); ++i) c += __builtin_popcountll(d.data()[i]);
              return c;
          },
          py::arg("det"));
    m.def("ctz",
          [](Array<u64> d) {
              for (i64 i = 0; i < (i64)d.size(); ++i)
                  if (d.data()[i]) return (i64)__builtin_ctzll(d.data()[i]) + 64 * i;
              return (i64)0;
          },
          py::arg("det"));
    m.def("add_hci", &add_hci<DOCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("eps") = 1.0e-5,
          py::arg("nthread") = -1, py::arg("selector") = py::none(), py::arg("pool") = 5.0, py::arg("nadd") = -1,
          py::arg("score_mode") = 0);
    m.def("add_hci", &add_hci<FullCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("eps") = 1.0e-5,
          py::arg("nthread") = -1, py::arg("selector") = py::none(), py::arg("pool") = 5.0, py::arg("nadd") = -1,
          py::arg("score_mode") = 0);
    m.def("add_hci", &add_hci<GenCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("eps") = 1.0e-5,
          py::arg("nthread") = -1, py::arg("selector") = py::none(), py::arg("pool") = 5.0, py::arg("nadd") = -1,
          py::arg("score_mode") = 0);
    m.def("compute_enpt2", &enpt2<DOCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("energy"),
          py::arg("eps") = 1.0e-5, py::arg("nthread") = -1);
    m.def("compute_enpt2", &enpt2<FullCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("energy"),
          py::arg("eps") = 1.0e-5, py::arg("nthread") = -1);
    m.def("compute_enpt2", &enpt2<GenCIWfn>, py::arg("ham"), py::arg("wfn"), py::arg("coeffs"), py::arg("energy"),
          py::arg("eps") = 1.0e-5, py::arg("nthread") = -1);
    m.def("compute_transition_rdms",
          [](DOCIWfn &a, DOCIWfn &b, Array<double> c1, Array<double> c2) { return trdms(a, b, c1, c2); },
          py::arg("wfn1"), py::arg("wfn2"), py::arg("coeffs1"), py::arg("coeffs2"));
    m.def("compute_transition_rdms",
          [](FullCIWfn &a, FullCIWfn &b, Array<double> c1, Array<double> c2) { return trdms(a, b, c1, c2); },
          py::arg("wfn1"), py::arg("wfn2"), py::arg("coeffs1"), py::arg("coeffs2"));
    m.def("compute_transition_rdms",
          [](GenCIWfn &a, GenCIWfn &b, Array<double> c1, Array<double> c2) { return trdms(a, b, c1, c2); },
          py::arg("wfn1"), py::arg("wfn2"), py::arg("coeffs1"), py::arg("coeffs2"));
    m.def("compute_rdms", [](DOCIWfn &w, Array<double> c) { return rdms(w, c); }, py::arg("wfn"), py::arg("coeffs"));
    m.def("compute_rdms", [](FullCIWfn &w, Array<double> c) { return rdms(w, c); }, py::arg("wfn"), py::arg("coeffs"));
    m.def("compute_rdms", [](GenCIWfn &w, Array<double> c) { return rdms(w, c); }, py::arg("wfn"), py::arg("coeffs"));
    m.def(
        "batch_solve",
        [](py::list problems, i64 n, double tol, const std::string &precision, i64 maxiter) {
            const i64 np = (i64)py::len(problems);
            if (np == 0) throw py::value_error("no problems");
            int prec;
            if (precision == "f64") prec = PYCI_PREC_F64;
            else if (precision == "mixed") prec = PYCI_PREC_MIXED;
            else throw py::value_error("precision must be 'f64' or 'mixed'");
            std::vector<i64> nb(np), nu(np), nd(np);
            std::vector<double> ec(np);
            std::vector<Array<double>> ones, twos;
            std::vector<const double *> op(np), tp(np);
            for (i64 p = 0; p < np; ++p) {
                py::tuple t = problems[p].cast<py::tuple>();
                if (t.size() != 3) throw py::value_error("each problem is (hamiltonian, nocc_up, nocc_dn)");
                SQuantOp &h = t[0].cast<SQuantOp &>();
                nb[p] = h.nbasis;
                nu[p] = t[1].cast<i64>();
                nd[p] = t[2].cast<i64>();
                ec[p] = h.ecore;
                ones.push_back(h.one_mo);
                twos.push_back(h.two_mo);
                op[p] = ones.back().data();
                tp[p] = twos.back().data();
            }
            Array<double> ev({np, n}), lat(np);
            Array<i64> it(np), ns(np);
            std::vector<int> conv(np);
            pyci_batch_stats st{};
            {
                py::gil_scoped_release nogil;
                check(pyci_batch_solve(context(), np, nb.data(), nu.data(), nd.data(), ec.data(), op.data(),
                                       tp.data(), n, tol, prec, maxiter, ev.mutable_data(), lat.mutable_data(),
                                       it.mutable_data(), ns.mutable_data(), conv.data(), &st));
            }
            py::dict info;
            info["latency_ms"] = lat;
            info["iterations"] = it;
            info["sigma_vectors"] = ns;
            py::list cv;
            for (int c : conv) cv.append(c != 0);
            info["converged"] = cv;
            info["graph_launches"] = st.graph_launches;
            info["rows_per_block"] = st.rows_per_block;
            info["setup_ms"] = st.setup_ms;
            info["total_ms"] = st.total_ms;
            info["single_path"] = st.single_path;
            return py::make_tuple(ev, info);
        },
        py::arg("problems"), py::arg("n") = 2, py::arg("tol") = 1.0e-6, py::arg("precision") = "mixed",
        py::arg("maxiter") = -1);
    py::class_<MLPSelector>(m, "mlp_selector")
        .def(py::init<i64, u64>(), py::arg("nfeat"), py::arg("seed") = 0)
        .def_readonly("nfeat", &MLPSelector::nfeat)
        .def_property("params", &MLPSelector::get_params, &MLPSelector::set_params)
        .def("predict", &MLPSelector::predict, py::arg("wfn"))
        .def("loss_grad", &MLPSelector::loss_grad, py::arg("wfn"), py::arg("target"))
        .def("train", &MLPSelector::train, py::arg("wfn"), py::arg("coeffs"), py::arg("steps") = 300,
             py::arg("batch") = 8192, py::arg("lr") = 1.0e-3, py::arg("cmin") = 1.0e-10, py::arg("seed") = 0);
    m.attr("MLP_NPARAM") = (i64)PYCI_MLP_NPARAM;
    m.def("device_count", [] {
        int c = 0;
        check(pyci_device_count(&c));
        return c;
    });
}
