// Host-side objects behind the pybind11 module: integrals, determinant
// containers and the engine context. The containers keep determinants in
// insertion order (as the reference's Wfn::dets) with a lazily built host
// index; the GPU engine (include/pyci_engine.h) gets a device mirror that is
// extended in place when only appends happened.
#pragma once

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pyci_engine.h"

namespace py = pybind11;

namespace ph {

using u64 = uint64_t;
using i64 = int64_t;

template <typename T>
using Array = py::array_t<T, py::array::c_style | py::array::forcecast>;

// ---------------------------------------------------------------------------
// Engine context and status -> Python exception mapping
// ---------------------------------------------------------------------------
struct EngineTypeError : std::exception {
    std::string m;
    explicit EngineTypeError(std::string s) : m(std::move(s)) {}
    const char *what() const noexcept override { return m.c_str(); }
};

[[noreturn]] void raise_status(int rc);
inline void check(int rc) {
    if (rc != PYCI_OK) raise_status(rc);
}
pyci_ctx *context();  // lazily created; throws RuntimeError without a GPU

// ---------------------------------------------------------------------------
// Combinatorics and bit helpers (the reference's common.cpp primitives)
// ---------------------------------------------------------------------------
i64 binomial(i64 n, i64 k);  // saturates at INT64_MAX
inline i64 nword_det(i64 n) { return n / 64 + ((n % 64) ? 1 : 0); }
void fill_hartreefock_det(i64 nocc, u64 *det);
void fill_det(i64 nocc, const i64 *occs, u64 *det);
void fill_occs(i64 nword, const u64 *det, i64 *occs);
void fill_virs(i64 nword, i64 nbasis, const u64 *det, i64 *virs);
void next_colex(i64 *indices);
i64 rank_colex(i64 nbasis, i64 nocc, const u64 *det);

// SpookyHash V2 128-bit (public-domain algorithm by Bob Jenkins) with the
// reference's seeds (pyci.h:123-128); used only for rank_det.
std::pair<u64, u64> spooky128(const void *msg, size_t len, u64 seed1, u64 seed2);

// ---------------------------------------------------------------------------
// Integrals (SQuantOp, squantop.cpp)
// ---------------------------------------------------------------------------
struct SQuantOp {
    i64 nbasis = 0;
    double ecore = 0.0;
    Array<double> one_mo, two_mo, h, v, w;
    std::shared_ptr<pyci_ham> dev;

    SQuantOp() = default;
    explicit SQuantOp(const std::string &filename);
    SQuantOp(double e, Array<double> mo1, Array<double> mo2);
    void derive();
    void to_file(const std::string &filename, i64 nelec, i64 ms2, double tol) const;
    pyci_ham *device();
};

// ---------------------------------------------------------------------------
// Open-addressing index over determinants stored in a contiguous array
// ---------------------------------------------------------------------------
struct HostIndex {
    std::vector<i64> slots;  // -1 empty, else det index
    i64 count = 0;
    void clear() {
        slots.clear();
        count = 0;
    }
    static u64 hash(const u64 *d, i64 nw);
    i64 find(const std::vector<u64> &dets, i64 nw, const u64 *d) const;
    void insert(const std::vector<u64> &dets, i64 nw, i64 idx);  // idx must be absent
    void rebuild(const std::vector<u64> &dets, i64 nw, i64 n);
};

// ---------------------------------------------------------------------------
// Wave functions (Wfn / OneSpinWfn / TwoSpinWfn, wfn.cpp, onespinwfn.cpp, twospinwfn.cpp)
// ---------------------------------------------------------------------------
struct Wfn {
    i64 nbasis = 0, nocc = 0, nocc_up = 0, nocc_dn = 0, nvir = 0, nvir_up = 0, nvir_dn = 0;
    i64 ndet = 0, nword = 0, nwdet = 0;  // nwdet: words per determinant (nword or 2*nword)
    i64 maxrank_up = 0, maxrank_dn = 0;
    int kind = PYCI_WFN_DOCI;
    // Host copy of the determinants: dets_ holds the first host_ndet of the ndet.
    // Determinants that add_hci appended on the device are copied here only when
    // host code first reads them (hdets(), det_ptr(), lookups), so that an
    // add_hci -> update -> solve loop keeps them on the device.
    mutable std::vector<u64> dets_;
    mutable i64 host_ndet = 0;
    mutable HostIndex index;
    mutable bool index_valid = true;
    bool full_colex = false;  // dets are exactly the full space in colex order
    // device mirror
    std::shared_ptr<pyci_wfn> dev;
    i64 dev_ndet = 0;

    Wfn() = default;
    Wfn(i64 nb, i64 nu, i64 nd, int kind_);
    virtual ~Wfn() = default;
    void init(i64 nb, i64 nu, i64 nd);

    i64 length() const { return ndet; }
    void squeeze() { dets_.shrink_to_fit(); }
    void reserve(i64 n) { dets_.reserve((size_t)(n * nwdet)); }
    void sync_host() const;              // copies device-only determinants to dets_
    const std::vector<u64> &hdets() const {
        sync_host();
        return dets_;
    }
    const u64 *det_ptr(i64 i) const { return &hdets()[(size_t)(i * nwdet)]; }
    i64 host_pending() const { return ndet - host_ndet; }  // determinants held only on the device
    void set_dets(std::vector<u64> &&buf, i64 n);  // replaces the set (host and device mirror)

    void ensure_index() const;
    i64 index_det(const u64 *d) const;
    i64 add_det(const u64 *d);          // -1 if present
    void append_unique(const u64 *d, i64 n);  // caller guarantees absence
    void reset();                        // drop all dets (add_all_dets)
    pyci_wfn *device();                  // synced device mirror
    void adopt_device_appends(i64 old_ndet);  // dets appended on the device
    void fill_all_dets();
    void to_file(const std::string &filename) const;
    void load_file(const std::string &filename);
    std::pair<u64, u64> rank_det(const u64 *d) const;
    i64 index_det_from_rank(std::pair<u64, u64> r) const;
};

struct OneSpinWfn : Wfn {
    using Wfn::Wfn;
    void add_hartreefock_det();
    i64 add_det_from_occs(const i64 *occs);
    void add_excited_dets(const u64 *rdet, i64 e);
    i64 py_add_excited_dets(i64 exc, py::object ref);
};

struct TwoSpinWfn : Wfn {
    using Wfn::Wfn;
    void add_hartreefock_det();
    i64 add_det_from_occs(const i64 *occs);
    void add_excited_dets(const u64 *rdet, i64 e_up, i64 e_dn);
    i64 py_add_excited_dets(i64 exc, py::object ref);
};

struct DOCIWfn : OneSpinWfn {
    DOCIWfn(i64 nb, i64 nu, i64 nd) : OneSpinWfn(nb, nu, nd, PYCI_WFN_DOCI) {
        if (nocc_up != nocc_dn) throw std::invalid_argument("nocc_up != nocc_dn");
    }
};

struct FullCIWfn : TwoSpinWfn {
    FullCIWfn(i64 nb, i64 nu, i64 nd) : TwoSpinWfn(nb, nu, nd, PYCI_WFN_FULLCI) {}
};

struct GenCIWfn : OneSpinWfn {
    GenCIWfn(i64 nb, i64 nu, i64 nd) : OneSpinWfn(nb, nu, nd, PYCI_WFN_GENCI) {
        if (nocc_dn) throw std::invalid_argument("nocc_dn != 0");
    }
};

// ---------------------------------------------------------------------------
// Sparse operator facade (SparseOp, sparseop.cpp)
// ---------------------------------------------------------------------------
struct SparseOp {
    std::shared_ptr<pyci_op> op;
    i64 nrow = 0, ncol = 0, size = 0;
    double ecore = 0.0;
    bool symmetric = true;
    py::tuple shape;
    std::string storage = "csr";

    SparseOp(SQuantOp &ham, Wfn &wfn, i64 nrow, i64 ncol, bool symm, const std::string &storage);
    void refresh();
    void update(SQuantOp &ham, Wfn &wfn);
    Array<double> matvec(Array<double> x) const;
    Array<double> matvec_out(Array<double> x, Array<double> out) const;
    double get_element(i64 i, i64 j) const;
    py::tuple solve(i64 n, py::object c0, i64 ncv, i64 maxiter, double tol, const std::string &precision) const;
    py::tuple export_csr() const;
    Array<double> diagonal() const;
};

}  // namespace ph
