// Determinant containers (wfn.cpp, onespinwfn.cpp, twospinwfn.cpp semantics).
#include <algorithm>
#include <cstring>
#include <fstream>

#include "host.hpp"

namespace ph {

static constexpr u64 SPOOKY_SEED1 = 0x23a23cf5033c3c81ULL, SPOOKY_SEED2 = 0xb3816f6a2c68e530ULL;

Wfn::Wfn(i64 nb, i64 nu, i64 nd, int kind_) : kind(kind_) { init(nb, nu, nd); }

// Wfn::init (wfn.cpp:51-70)
void Wfn::init(i64 nb, i64 nu, i64 nd) {
    if (nd < 0) throw std::domain_error("nocc_dn is < 0");
    if (nu < nd) throw std::domain_error("nocc_up is < nocc_dn");
    if (nb < nu) throw std::domain_error("nbasis is < nocc_up");
    nbasis = nb;
    nocc = nu + nd;
    nocc_up = nu;
    nocc_dn = nd;
    nvir = nb * 2 - nu - nd;
    nvir_up = nb - nu;
    nvir_dn = nb - nd;
    ndet = 0;
    host_ndet = 0;
    nword = nword_det(nb);
    nwdet = (kind == PYCI_WFN_FULLCI) ? 2 * nword : nword;
    maxrank_up = binomial(nb, nu);
    maxrank_dn = binomial(nb, nd);
    dets_.clear();
    index.clear();
    index_valid = true;
    full_colex = false;
    dev.reset();
    dev_ndet = 0;
}

// Copies the determinants that exist only on the device (appended there by
// add_hci) to the host copy and its index.
void Wfn::sync_host() const {
    if (host_ndet >= ndet) return;
    const i64 n = ndet - host_ndet;
    const size_t at = dets_.size();
    dets_.resize(at + (size_t)(n * nwdet));
    check(pyci_wfn_get_dets(dev.get(), host_ndet, ndet, dets_.data() + at));
    if (index_valid)
        for (i64 i = host_ndet; i < ndet; ++i) index.insert(dets_, nwdet, i);
    host_ndet = ndet;
}

void Wfn::set_dets(std::vector<u64> &&buf, i64 n) {
    reset();
    dets_.swap(buf);
    ndet = host_ndet = n;
    index_valid = false;  // built only when a host lookup needs it
}

void Wfn::ensure_index() const {
    sync_host();
    if (!index_valid) {
        index.rebuild(dets_, nwdet, ndet);
        index_valid = true;
    }
}

i64 Wfn::index_det(const u64 *d) const {
    ensure_index();
    return index.find(dets_, nwdet, d);
}

i64 Wfn::add_det(const u64 *d) {
    if (index_det(d) >= 0) return -1;
    dets_.insert(dets_.end(), d, d + nwdet);
    index.insert(dets_, nwdet, ndet);
    full_colex = false;
    host_ndet = ndet + 1;
    return ndet++;
}

void Wfn::append_unique(const u64 *d, i64 n) {
    if (n <= 0) return;
    sync_host();
    dets_.insert(dets_.end(), d, d + n * nwdet);
    if (index_valid) {
        for (i64 i = 0; i < n; ++i) index.insert(dets_, nwdet, ndet + i);
    }
    ndet += n;
    host_ndet = ndet;
    full_colex = false;
}

void Wfn::reset() {
    dets_.clear();
    ndet = 0;
    host_ndet = 0;
    index.clear();
    index_valid = true;
    full_colex = false;
    dev.reset();
    dev_ndet = 0;
}

// Device mirror: created on first use; extended with pyci_wfn_append when the
// host only appended since the last sync; recreated otherwise.
pyci_wfn *Wfn::device() {
    if (dev && dev_ndet == ndet) return dev.get();
    if (dev && dev_ndet < ndet) {
        check(pyci_wfn_append(dev.get(), ndet - dev_ndet, det_ptr(dev_ndet)));
        dev_ndet = ndet;
        return dev.get();
    }
    if (nbasis > 64) throw py::value_error("the GPU engine supports nbasis <= 64");
    pyci_ctx *ctx = context();
    pyci_wfn *w = nullptr;
    check(pyci_wfn_create(ctx, kind, nbasis, nocc_up, nocc_dn, ndet, hdets().data(), &w));
    dev = std::shared_ptr<pyci_wfn>(w, pyci_wfn_destroy);
    dev_ndet = ndet;
    return w;
}

// Determinants appended on the device (add_hci) join the set; the host copy
// follows only when host code reads them (sync_host).
void Wfn::adopt_device_appends(i64 old_ndet) {
    i64 nd = 0;
    check(pyci_wfn_ndet(dev.get(), &nd));
    if (nd <= old_ndet) return;
    ndet = nd;
    dev_ndet = nd;
    full_colex = false;
}

// add_all_dets (onespinwfn.cpp:173-217, twospinwfn.cpp:223-245): the full space
// in colex order, generated on the GPU.
void Wfn::fill_all_dets() {
    const i64 big = std::numeric_limits<i64>::max();
    if (maxrank_up == big || (kind == PYCI_WFN_FULLCI && (maxrank_dn == big || maxrank_up > big / std::max<i64>(maxrank_dn, 1))))
        throw std::domain_error("cannot generate > 2 ** 63 determinants");
    const i64 n = (kind == PYCI_WFN_FULLCI) ? maxrank_up * maxrank_dn : maxrank_up;
    std::vector<u64> buf((size_t)(n * nwdet), 0ULL);
    if (n > 0) check(pyci_fill_all_dets(context(), kind, nbasis, nocc_up, nocc_dn, buf.data()));
    set_dets(std::move(buf), n);
    full_colex = true;
}

// Binary format (onespinwfn.cpp:92-104, twospinwfn.cpp:95-107):
// int64 ndet, nbasis, nocc_up, nocc_dn, then the determinant words.
void Wfn::to_file(const std::string &filename) const {
    std::ofstream f(filename, std::ios::out | std::ios::binary);
    i64 hdr[4] = {ndet, nbasis, nocc_up, nocc_dn};
    bool ok = (bool)f.write((const char *)hdr, sizeof(hdr));
    if (ok && ndet) ok = (bool)f.write((const char *)hdets().data(), (std::streamsize)(sizeof(u64) * ndet * nwdet));
    f.close();
    if (!ok) throw std::ios_base::failure("error writing file");
}

void Wfn::load_file(const std::string &filename) {
    std::ifstream f(filename, std::ios::in | std::ios::binary);
    i64 hdr[4];
    if (!f.read((char *)hdr, sizeof(hdr))) throw std::ios_base::failure("error in file");
    init(hdr[1], hdr[2], hdr[3]);
    std::vector<u64> buf((size_t)(hdr[0] * nwdet));
    if (hdr[0] && !f.read((char *)buf.data(), (std::streamsize)(sizeof(u64) * buf.size())))
        throw std::ios_base::failure("error in file");
    set_dets(std::move(buf), hdr[0]);
}

std::pair<u64, u64> Wfn::rank_det(const u64 *d) const {
    return spooky128(d, sizeof(u64) * (size_t)nwdet, SPOOKY_SEED1, SPOOKY_SEED2);
}

i64 Wfn::index_det_from_rank(std::pair<u64, u64> r) const {
    for (i64 i = 0; i < ndet; ++i)
        if (rank_det(det_ptr(i)) == r) return i;
    return -1;
}

// ---------------------------------------------------------------------------
// One-spin
// ---------------------------------------------------------------------------
void OneSpinWfn::add_hartreefock_det() {
    std::vector<u64> d((size_t)nword, 0ULL);
    fill_hartreefock_det(nocc_up, d.data());
    add_det(d.data());
}

i64 OneSpinWfn::add_det_from_occs(const i64 *occs) {
    std::vector<u64> d((size_t)nword, 0ULL);
    fill_det(nocc_up, occs, d.data());
    return add_det(d.data());
}

// add_excited_dets (onespinwfn.cpp:219-244): virtual combinations (outer, colex)
// x occupied combinations (inner, colex).
void OneSpinWfn::add_excited_dets(const u64 *rdet, i64 e) {
    const i64 no = binomial(nocc_up, e), nv = binomial(nvir_up, e);
    if (no == 0 || nv == 0) return;
    std::vector<u64> d((size_t)nword);
    std::vector<i64> occs((size_t)nocc_up + 1), virs((size_t)nvir_up + 1), oi((size_t)e + 1), vi((size_t)e + 1);
    fill_occs(nword, rdet, occs.data());
    fill_virs(nword, nbasis, rdet, virs.data());
    for (i64 k = 0; k < e; ++k) vi[(size_t)k] = k;
    vi[(size_t)e] = nvir_up + 1;
    for (i64 i = 0; i < nv; ++i) {
        for (i64 k = 0; k < e; ++k) oi[(size_t)k] = k;
        oi[(size_t)e] = nocc_up + 1;
        for (i64 j = 0; j < no; ++j) {
            std::memcpy(d.data(), rdet, sizeof(u64) * nword);
            for (i64 k = 0; k < e; ++k) {
                i64 o = occs[(size_t)oi[(size_t)k]], v = virs[(size_t)vi[(size_t)k]];
                d[(size_t)(o / 64)] &= ~(1ULL << (o % 64));
                d[(size_t)(v / 64)] |= 1ULL << (v % 64);
            }
            add_det(d.data());
            if (e > 0) next_colex(oi.data());
        }
        if (e > 0) next_colex(vi.data());
    }
}

static std::vector<u64> ref_or_hf(const Wfn &w, py::object ref, bool two_spin) {
    std::vector<u64> r((size_t)w.nwdet, 0ULL);
    if (ref.is_none()) {
        fill_hartreefock_det(w.nocc_up, r.data());
        if (two_spin) fill_hartreefock_det(w.nocc_dn, r.data() + w.nword);
    } else {
        auto a = ref.cast<Array<u64>>();
        if ((i64)a.size() != w.nwdet) throw py::value_error("ref has the wrong number of words");
        std::memcpy(r.data(), a.data(), sizeof(u64) * (size_t)w.nwdet);
    }
    return r;
}

i64 OneSpinWfn::py_add_excited_dets(i64 exc, py::object ref) {
    auto r = ref_or_hf(*this, ref, false);
    i64 old = ndet;
    add_excited_dets(r.data(), exc);
    return ndet - old;
}

// ---------------------------------------------------------------------------
// Two-spin
// ---------------------------------------------------------------------------
void TwoSpinWfn::add_hartreefock_det() {
    std::vector<u64> d((size_t)nwdet, 0ULL);
    fill_hartreefock_det(nocc_up, d.data());
    fill_hartreefock_det(nocc_dn, d.data() + nword);
    add_det(d.data());
}

i64 TwoSpinWfn::add_det_from_occs(const i64 *occs) {
    std::vector<u64> d((size_t)nwdet, 0ULL);
    fill_det(nocc_up, occs, d.data());
    fill_det(nocc_dn, occs + nocc_up, d.data() + nword);
    return add_det(d.data());
}

// twospinwfn.cpp:247-265: alpha excitations (outer) x beta excitations (inner).
void TwoSpinWfn::add_excited_dets(const u64 *rdet, i64 e_up, i64 e_dn) {
    if (e_up == 0 && e_dn == 0) {
        add_det(rdet);
        return;
    }
    OneSpinWfn wu(nbasis, nocc_up, nocc_up, PYCI_WFN_GENCI), wd(nbasis, nocc_dn, nocc_dn, PYCI_WFN_GENCI);
    wu.add_excited_dets(rdet, e_up);
    wd.add_excited_dets(rdet + nword, e_dn);
    std::vector<u64> d((size_t)nwdet);
    for (i64 i = 0; i < wu.ndet; ++i) {
        std::memcpy(d.data(), wu.det_ptr(i), sizeof(u64) * nword);
        for (i64 j = 0; j < wd.ndet; ++j) {
            std::memcpy(d.data() + nword, wd.det_ptr(j), sizeof(u64) * nword);
            add_det(d.data());
        }
    }
}

// twospinwfn.cpp:325-342
i64 TwoSpinWfn::py_add_excited_dets(i64 exc, py::object ref) {
    auto r = ref_or_hf(*this, ref, true);
    i64 old = ndet;
    i64 maxup = std::min(nocc_up, nvir_up), maxdn = std::min(nocc_dn, nvir_dn);
    i64 a = std::min(exc, maxup), b = exc - a;
    while (a >= 0 && b <= maxdn) add_excited_dets(r.data(), a--, b++);
    return ndet - old;
}

}  // namespace ph
