// extern "C" entry points of include/pyci_engine.h.
#include <cstring>
#include <memory>
#include <string>

#include "internal.hpp"

namespace pe {

static thread_local std::string g_last_error = "";

void set_error(const std::string &msg) { g_last_error = msg; }

// declared in other translation units
void wfn_index_append(pyci_wfn *w, i64 first);
void wfn_lookup_batch(const pyci_wfn *w, const u64 *d_q, i64 n, i64 *d_out);
void export_csr(pyci_op *op, double *data, i64 *indices, i64 *indptr);
double get_element(pyci_op *op, i64 i, i64 j);
void op_bench_sigma(pyci_op *op, i64 nvec, i64 nrep, int fp32, float *avg_ms, double *bytes);

template <class F>
static int guarded(F &&f) {
    try {
        f();
        return PYCI_OK;
    } catch (const EngineError &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return PYCI_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return PYCI_EHIP;
    } catch (...) {
        set_error("unknown error");
        return PYCI_EHIP;
    }
}

static void set_device(pyci_ctx *ctx) { PE_HIP(hipSetDevice(ctx->device)); }

// swap_nb of an operator over wfn's rows (see pyci_op::swap_nb)
static i64 swap_partner_nb(const pyci_op *op, const pyci_wfn *wfn) {
    const bool full = wfn->kind == PYCI_WFN_FULLCI && wfn->colex && wfn->nup == wfn->ndn &&
                      op->nrow == wfn->ndet && op->ncol == wfn->ndet && op->symmetric;
    return full ? wfn->nsd : 0;
}

static void validate_kind(int kind, i64 nb, i64 nup, i64 ndn) {
    PE_REQUIRE(kind == PYCI_WFN_DOCI || kind == PYCI_WFN_FULLCI || kind == PYCI_WFN_GENCI, PYCI_EINVAL,
               "unknown wave function kind");
    PE_REQUIRE(ndn >= 0 && nup >= ndn && nb >= nup, PYCI_EINVAL, "invalid occupation numbers");
}

// Builds rows [0, nrow) with the given ncol into a fresh operator payload.
static void op_build_all(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn) {
    DevBuf<i64> rp, lc;
    DevBuf<int32_t> cols;
    DevBuf<double> vals, dg;
    i64 nnz = 0, low = 0;
    build_rows(op, ham, wfn, 0, op->nrow, op->ncol, rp, cols, vals, lc, dg, nnz, low);
    op->rowptr = std::move(rp);
    op->lowcnt = std::move(lc);
    op->cols = std::move(cols);
    op->vals = std::move(vals);
    op->diag = std::move(dg);
    op->vals32.release();
    op->nnz = nnz;
    op->size_ref = op->symmetric ? low : nnz;
    op->segments.assign(1, {op->nrow, op->ncol});
}

__global__ void k_add_offset(i64 *a, i64 n, i64 off) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += off;
}

}  // namespace pe

using namespace pe;

extern "C" {

const char *pyci_last_error(void) { return g_last_error.c_str(); }

int pyci_device_count(int *count) {
    return guarded([&] {
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

int pyci_ctx_create(int device, pyci_ctx **out) {
    return guarded([&] {
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        PE_REQUIRE(e == hipSuccess && c > 0, PYCI_EHIP, "no GPU available (hipGetDeviceCount found none)");
        PE_REQUIRE(device >= 0 && device < c, PYCI_EINVAL, "device index out of range");
        auto *ctx = new pyci_ctx();
        try {
            ctx->device = device;
            PE_HIP(hipSetDevice(device));
            PE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            PE_HIP(hipEventCreate(&ctx->ev0));
            PE_HIP(hipEventCreate(&ctx->ev1));
            std::vector<i64> b(65 * 65);
            for (int n = 0; n <= 64; ++n)
                for (int k = 0; k <= 64; ++k) b[(size_t)(n * 65 + k)] = binom_host(n, k);
            ctx->binom.alloc(65 * 65);
            PE_HIP(hipMemcpy(ctx->binom.p, b.data(), sizeof(i64) * 65 * 65, hipMemcpyHostToDevice));
            ctx->noalloc_ws = noalloc_scratch(device);
        } catch (...) {
            pyci_ctx_destroy(ctx);
            throw;
        }
        *out = ctx;
    });
}

void pyci_ctx_destroy(pyci_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    ctx->binom.release();
    ctx->mf_ws.release();
    ctx->mf_ws2.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
}

int pyci_ctx_synchronize(pyci_ctx *ctx) {
    return guarded([&] { PE_HIP(hipStreamSynchronize(ctx->stream)); });
}

int pyci_ctx_timer_start(pyci_ctx *ctx) {
    return guarded([&] { PE_HIP(hipEventRecord(ctx->ev0, ctx->stream)); });
}

int pyci_ctx_timer_stop(pyci_ctx *ctx, float *ms) {
    return guarded([&] {
        PE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PE_HIP(hipEventSynchronize(ctx->ev1));
        PE_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    });
}

int pyci_ham_create(pyci_ctx *ctx, int64_t nbasis, double ecore, const double *one_mo, const double *two_mo,
                    pyci_ham **out) {
    return guarded([&] {
        PE_REQUIRE(nbasis > 0 && nbasis <= 64, PYCI_EUNSUPPORTED, "nbasis must be in 1..64 for the GPU engine");
        set_device(ctx);
        auto *h = new pyci_ham();
        try {
            h->ctx = ctx;
            h->nbasis = nbasis;
            h->ecore = ecore;
            const size_t n2 = (size_t)(nbasis * nbasis), n4 = n2 * n2;
            h->one_mo.alloc(n2);
            h->two_mo.alloc(n4);
            PE_HIP(hipMemcpyAsync(h->one_mo.p, one_mo, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
            PE_HIP(hipMemcpyAsync(h->two_mo.p, two_mo, sizeof(double) * n4, hipMemcpyHostToDevice, ctx->stream));
            PE_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void pyci_ham_destroy(pyci_ham *h) { delete h; }

int pyci_wfn_create(pyci_ctx *ctx, int kind, int64_t nbasis, int64_t nup, int64_t ndn, int64_t ndet,
                    const uint64_t *dets, pyci_wfn **out) {
    return guarded([&] {
        validate_kind(kind, nbasis, nup, ndn);
        PE_REQUIRE(nbasis <= 64, PYCI_EUNSUPPORTED, "the GPU engine supports nbasis <= 64 (one word per string)");
        PE_REQUIRE(ndet >= 0, PYCI_EINVAL, "ndet < 0");
        set_device(ctx);
        auto *w = new pyci_wfn();
        try {
            w->ctx = ctx;
            w->kind = kind;
            w->nbasis = nbasis;
            w->nup = nup;
            w->ndn = ndn;
            w->nkey = (kind == PYCI_WFN_FULLCI) ? 2 : 1;
            w->nsd = (kind == PYCI_WFN_FULLCI) ? binom_host(nbasis, ndn) : 1;
            w->ndet = ndet;
            w->dets.alloc((size_t)(std::max<i64>(ndet, 16) * w->nkey));
            if (ndet > 0)
                PE_HIP(hipMemcpyAsync(w->dets.p, dets, sizeof(u64) * ndet * w->nkey, hipMemcpyHostToDevice,
                                      ctx->stream));
            wfn_build_index(w);
            PE_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete w;
            throw;
        }
        *out = w;
    });
}

int pyci_wfn_append(pyci_wfn *w, int64_t n, const uint64_t *dets) {
    return guarded([&] {
        PE_REQUIRE(n >= 0, PYCI_EINVAL, "n < 0");
        if (n == 0) return;
        set_device(w->ctx);
        hipStream_t s = w->ctx->stream;
        const i64 need = (w->ndet + n) * w->nkey;
        if ((size_t)need > w->dets.n) {
            DevBuf<u64> nb((size_t)(need + need / 2));
            if (w->ndet > 0)
                PE_HIP(hipMemcpyAsync(nb.p, w->dets.p, sizeof(u64) * w->ndet * w->nkey, hipMemcpyDeviceToDevice, s));
            PE_HIP(hipStreamSynchronize(s));
            w->dets = std::move(nb);
        }
        PE_HIP(hipMemcpyAsync(w->dets.p + w->ndet * w->nkey, dets, sizeof(u64) * n * w->nkey, hipMemcpyHostToDevice,
                              s));
        const i64 first = w->ndet;
        w->ndet += n;
        wfn_index_append(w, first);
        PE_HIP(hipStreamSynchronize(s));
    });
}

int pyci_wfn_ndet(const pyci_wfn *w, int64_t *ndet) {
    return guarded([&] { *ndet = w->ndet; });
}

int pyci_wfn_get_dets(const pyci_wfn *w, int64_t start, int64_t end, uint64_t *out) {
    return guarded([&] {
        PE_REQUIRE(start >= 0 && end <= w->ndet && start <= end, PYCI_EINVAL, "range out of bounds");
        if (end == start) return;
        PE_HIP(hipMemcpyAsync(out, w->dets.p + start * w->nkey, sizeof(u64) * (end - start) * w->nkey,
                              hipMemcpyDeviceToHost, w->ctx->stream));
        PE_HIP(hipStreamSynchronize(w->ctx->stream));
    });
}

int pyci_wfn_index(const pyci_wfn *w, int64_t n, const uint64_t *dets, int64_t *index) {
    return guarded([&] {
        if (n <= 0) return;
        set_device(w->ctx);
        hipStream_t s = w->ctx->stream;
        DevBuf<u64> q((size_t)(n * w->nkey));
        DevBuf<i64> o((size_t)n);
        PE_HIP(hipMemcpyAsync(q.p, dets, sizeof(u64) * n * w->nkey, hipMemcpyHostToDevice, s));
        wfn_lookup_batch(w, q.p, n, o.p);
        PE_HIP(hipMemcpyAsync(index, o.p, sizeof(i64) * n, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
    });
}

void pyci_wfn_destroy(pyci_wfn *w) { delete w; }

int pyci_fill_all_dets(pyci_ctx *ctx, int kind, int64_t nbasis, int64_t nup, int64_t ndn, uint64_t *out) {
    return guarded([&] {
        validate_kind(kind, nbasis, nup, ndn);
        set_device(ctx);
        const int nword = (int)((nbasis + 63) / 64);
        i64 n = binom_host(nbasis, nup) * (kind == PYCI_WFN_FULLCI ? binom_host(nbasis, ndn) : 1);
        PE_REQUIRE(n >= 0 && n < ((i64)1 << 40), PYCI_EINVAL, "determinant space too large");
        const i64 words = n * nword * (kind == PYCI_WFN_FULLCI ? 2 : 1);
        if (words == 0) return;
        DevBuf<u64> d((size_t)words);
        fill_all_dets_device(ctx, kind, nbasis, nup, ndn, d.p, nword);
        PE_HIP(hipMemcpyAsync(out, d.p, sizeof(u64) * words, hipMemcpyDeviceToHost, ctx->stream));
        PE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pyci_op_create(pyci_ham *ham, pyci_wfn *wfn, int64_t nrow, int64_t ncol, int symmetric, pyci_op **out) {
    return pyci_op_create_ex(ham, wfn, nrow, ncol, symmetric, PYCI_STORAGE_CSR, out);
}

int pyci_op_create_ex(pyci_ham *ham, pyci_wfn *wfn, int64_t nrow, int64_t ncol, int symmetric, int storage,
                      pyci_op **out) {
    return guarded([&] {
        PE_REQUIRE(storage == PYCI_STORAGE_CSR || storage == PYCI_STORAGE_DIRECT || storage == PYCI_STORAGE_AUTO,
                   PYCI_EINVAL, "unknown storage");
        PE_REQUIRE(ham->ctx == wfn->ctx, PYCI_EINVAL, "ham and wfn belong to different contexts");
        const i64 nbexp = (wfn->kind == PYCI_WFN_GENCI) ? wfn->nbasis : wfn->nbasis;
        PE_REQUIRE(ham->nbasis == nbexp, PYCI_EINVAL, "operator and wave function have different nbasis");
        set_device(wfn->ctx);
        auto *op = new pyci_op();
        try {
            op->ctx = wfn->ctx;
            op->kind = wfn->kind;
            op->nrow = (nrow > -1) ? nrow : wfn->ndet;
            op->ncol = (ncol > -1) ? ncol : wfn->ndet;
            PE_REQUIRE(op->nrow <= wfn->ndet, PYCI_EINVAL, "nrow exceeds the number of determinants");
            PE_REQUIRE(op->ncol <= wfn->ndet, PYCI_EINVAL, "ncol exceeds the number of determinants");
            PE_REQUIRE(op->ncol < ((i64)1 << 31), PYCI_EUNSUPPORTED, "ncol must be < 2^31");
            op->symmetric = symmetric ? 1 : 0;
            op->ecore = ham->ecore;
            const bool direct = storage == PYCI_STORAGE_DIRECT ||
                                (storage == PYCI_STORAGE_AUTO &&
                                 direct_applicable(ham, wfn, op->nrow, op->ncol, op->symmetric));
            if (direct) {
                op->storage = PYCI_STORAGE_DIRECT;
                direct_setup(op, ham, wfn);
            } else {
                op_build_all(op, ham, wfn);
            }
            op->swap_nb = swap_partner_nb(op, wfn);
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
    });
}

int pyci_op_storage(const pyci_op *op, int *storage) {
    return guarded([&] { *storage = op->storage; });
}

int pyci_op_update(pyci_op *op, pyci_ham *ham, pyci_wfn *wfn) {
    return guarded([&] {
        set_device(op->ctx);
        PE_REQUIRE(wfn->ndet < ((i64)1 << 31), PYCI_EUNSUPPORTED, "ncol must be < 2^31");
        if (op->storage == PYCI_STORAGE_DIRECT) {
            // unchanged full space: nothing to do; otherwise continue as a CSR operator
            if (wfn->ndet == op->nrow && direct_applicable(ham, wfn, wfn->ndet, wfn->ndet, op->symmetric)) return;
            op->storage = PYCI_STORAGE_CSR;
            op->direct.reset();
            op->nrow = wfn->ndet;
            op->ncol = wfn->ndet;
            op_build_all(op, ham, wfn);
            op->swap_nb = swap_partner_nb(op, wfn);
            return;
        }
        op->swap_nb = 0;
        const i64 old = op->nrow;
        if (op->symmetric) {
            // The lower-triangle export of rows [0, old) is unchanged by new columns
            // (they are all > row); the stored full pattern gains upper entries, so
            // the operator is rebuilt over the new square shape.
            op->nrow = wfn->ndet;
            op->ncol = wfn->ndet;
            op_build_all(op, ham, wfn);
            op->swap_nb = swap_partner_nb(op, wfn);
            return;
        }
        // Non-symmetric: old rows keep their column bound, new rows use ndet
        // (update(ham, wfn, ndet, ndet, nrow), sparseop.cpp:175-201).
        const i64 newrows = wfn->ndet;
        if (newrows <= old) {
            op->ncol = wfn->ndet;
            return;
        }
        hipStream_t s = op->ctx->stream;
        DevBuf<i64> rp, lc;
        DevBuf<int32_t> cols;
        DevBuf<double> vals, dg;
        i64 nnz = 0, low = 0;
        build_rows(op, ham, wfn, old, newrows, wfn->ndet, rp, cols, vals, lc, dg, nnz, low);
        const i64 tot = op->nnz + nnz;
        DevBuf<i64> nrp((size_t)(newrows + 1)), nlc((size_t)(newrows + 1));
        DevBuf<int32_t> ncols((size_t)tot);
        DevBuf<double> nvals((size_t)tot), ndg((size_t)newrows);
        PE_HIP(hipMemcpyAsync(nrp.p, op->rowptr.p, sizeof(i64) * old, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipMemcpyAsync(nrp.p + old, rp.p, sizeof(i64) * (newrows - old + 1), hipMemcpyDeviceToDevice, s));
        k_add_offset<<<grid_for(newrows - old + 1, 256), 256, 0, s>>>(nrp.p + old, newrows - old + 1, op->nnz);
        PE_LAUNCH_CHECK();
        PE_HIP(hipMemcpyAsync(nlc.p, op->lowcnt.p, sizeof(i64) * old, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipMemcpyAsync(nlc.p + old, lc.p, sizeof(i64) * (newrows - old + 1), hipMemcpyDeviceToDevice, s));
        if (op->nnz) {
            PE_HIP(hipMemcpyAsync(ncols.p, op->cols.p, sizeof(int32_t) * op->nnz, hipMemcpyDeviceToDevice, s));
            PE_HIP(hipMemcpyAsync(nvals.p, op->vals.p, sizeof(double) * op->nnz, hipMemcpyDeviceToDevice, s));
        }
        if (nnz) {
            PE_HIP(hipMemcpyAsync(ncols.p + op->nnz, cols.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, s));
            PE_HIP(hipMemcpyAsync(nvals.p + op->nnz, vals.p, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
        }
        if (old) PE_HIP(hipMemcpyAsync(ndg.p, op->diag.p, sizeof(double) * old, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipMemcpyAsync(ndg.p + old, dg.p, sizeof(double) * (newrows - old), hipMemcpyDeviceToDevice, s));
        PE_HIP(hipStreamSynchronize(s));
        op->rowptr = std::move(nrp);
        op->lowcnt = std::move(nlc);
        op->cols = std::move(ncols);
        op->vals = std::move(nvals);
        op->diag = std::move(ndg);
        op->vals32.release();
        op->nnz = tot;
        op->size_ref = tot;
        op->nrow = newrows;
        op->ncol = wfn->ndet;
        op->segments.push_back({newrows, wfn->ndet});
    });
}

int pyci_op_info(const pyci_op *op, int64_t *nrow, int64_t *ncol, int64_t *size, int64_t *nnz_stored, int *symmetric,
                 double *ecore) {
    return guarded([&] {
        if (nrow) *nrow = op->nrow;
        if (ncol) *ncol = op->ncol;
        if (size) *size = op->size_ref;
        if (nnz_stored) *nnz_stored = op->nnz;
        if (symmetric) *symmetric = op->symmetric;
        if (ecore) *ecore = op->ecore;
    });
}

int pyci_op_export(pyci_op *op, double *data, int64_t *indices, int64_t *indptr) {
    return guarded([&] {
        set_device(op->ctx);
        direct_ensure_csr(op);
        export_csr(op, data, indices, indptr);
    });
}

int pyci_op_get_element(pyci_op *op, int64_t i, int64_t j, double *value) {
    return guarded([&] {
        PE_REQUIRE(i >= 0 && i < op->nrow && j >= 0, PYCI_EINVAL, "index out of range");
        set_device(op->ctx);
        if (op->symmetric && j > i) {
            *value = 0.0;
            return;
        }
        direct_ensure_csr(op);
        *value = get_element(op, i, j);
    });
}

int pyci_op_diagonal(pyci_op *op, double *diag) {
    return guarded([&] {
        set_device(op->ctx);
        if (op->nrow == 0) return;
        PE_HIP(hipMemcpyAsync(diag, op->diag.p, sizeof(double) * op->nrow, hipMemcpyDeviceToHost, op->ctx->stream));
        PE_HIP(hipStreamSynchronize(op->ctx->stream));
    });
}

static void check_matvec_shape(const pyci_op *op) {
    PE_REQUIRE(!op->symmetric || op->nrow == op->ncol, PYCI_EINVAL,
               "symmetric operator products need nrow == ncol");
}

int pyci_op_matvec_device(pyci_op *op, int64_t nvec, const double *x, double *y) {
    return guarded([&] {
        PE_REQUIRE(nvec >= 1, PYCI_EINVAL, "nvec < 1");
        check_matvec_shape(op);
        set_device(op->ctx);
        sigma_f64(op, (int)nvec, x, y, op->ctx->stream);
    });
}

int pyci_op_matvec_device_noalloc(pyci_op *op, int64_t nvec, const double *x, double *y) {
    return guarded([&] {
        PE_REQUIRE(op && x && y, PYCI_EINVAL, "null argument");
        PE_REQUIRE(nvec >= 1, PYCI_EINVAL, "nvec < 1");
        check_matvec_shape(op);
        set_device(op->ctx);
        if (op->storage == PYCI_STORAGE_DIRECT)
            direct_sigma_noalloc(op, (int)nvec, x, y, op->ctx->noalloc_ws, op->ctx->stream);
        else
            sigma_f64(op, (int)nvec, x, y, op->ctx->stream);  // the CSR product uses no workspace
    });
}

int pyci_op_matvec(pyci_op *op, int64_t nvec, const double *x, double *y) {
    return guarded([&] {
        PE_REQUIRE(nvec >= 1, PYCI_EINVAL, "nvec < 1");
        check_matvec_shape(op);
        set_device(op->ctx);
        hipStream_t s = op->ctx->stream;
        const i64 nc = op->ncol, nr = op->nrow;
        DevBuf<double> dx((size_t)(nc * nvec)), dy((size_t)(nr * nvec));
        if (nvec == 1) {
            PE_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * nc, hipMemcpyHostToDevice, s));
            sigma_f64(op, 1, dx.p, dy.p, s);
            PE_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * nr, hipMemcpyDeviceToHost, s));
        } else {
            // host layout: one vector per row; device layout: one determinant per row
            std::vector<double> t((size_t)(nc * nvec)), u((size_t)(nr * nvec));
            for (i64 v = 0; v < nvec; ++v)
                for (i64 c = 0; c < nc; ++c) t[(size_t)(c * nvec + v)] = x[v * nc + c];
            PE_HIP(hipMemcpyAsync(dx.p, t.data(), sizeof(double) * nc * nvec, hipMemcpyHostToDevice, s));
            sigma_f64(op, (int)nvec, dx.p, dy.p, s);
            PE_HIP(hipMemcpyAsync(u.data(), dy.p, sizeof(double) * nr * nvec, hipMemcpyDeviceToHost, s));
            PE_HIP(hipStreamSynchronize(s));
            for (i64 v = 0; v < nvec; ++v)
                for (i64 r = 0; r < nr; ++r) y[v * nr + r] = u[(size_t)(r * nvec + v)];
        }
        PE_HIP(hipStreamSynchronize(s));
    });
}

int pyci_op_bench_sigma(pyci_op *op, int64_t nvec, int64_t nrep, int fp32, float *avg_ms, double *bytes) {
    return guarded([&] {
        check_matvec_shape(op);
        set_device(op->ctx);
        op_bench_sigma(op, nvec, nrep, fp32, avg_ms, bytes);
    });
}

int pyci_op_bench_sigma_parts(pyci_op *op, int64_t nvec, int64_t nrep, int fp32, int *nparts, float *avg_ms,
                              double *bytes, double *flops) {
    return guarded([&] {
        check_matvec_shape(op);
        PE_REQUIRE(nvec >= 1 && nrep >= 1, PYCI_EINVAL, "nvec and nrep must be >= 1");
        set_device(op->ctx);
        if (op->storage == PYCI_STORAGE_DIRECT) {
            *nparts = 2;
            direct_bench_parts(op, nvec, nrep, fp32, avg_ms, bytes, flops);
        } else {
            *nparts = 1;
            op_bench_sigma(op, nvec, nrep, fp32, avg_ms, bytes);
            flops[0] = 2.0 * (double)op->nnz * (double)nvec;
        }
    });
}

int pyci_op_solve(pyci_op *op, int64_t n, const double *c0, int64_t ncv, int64_t maxiter, double tol, int precision,
                  double *evals, double *evecs, pyci_solve_stats *stats) {
    return guarded([&] {
        const i64 nr = op->nrow;
        PE_REQUIRE(n >= 1, PYCI_EINVAL, "n must be >= 1");
        PE_REQUIRE(!((nr > 1 && n >= nr) || (nr == 1 && n > 1)), PYCI_EINVAL,
                   "cannot find >=n eigenpairs for sparse operator with n rows");
        if (op->nrow != op->ncol) throw EngineError(PYCI_ETYPE, "Can only solve sparse symmetric matrix operators");
        PE_REQUIRE(nr >= 1, PYCI_EINVAL, "empty operator");
        set_device(op->ctx);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (nr == 1) {
            // the diagonal is written by the build kernels on the context stream
            double d = 0.0;
            PE_HIP(hipMemcpyAsync(&d, op->diag.p, sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
            PE_HIP(hipStreamSynchronize(op->ctx->stream));
            evals[0] = d + op->ecore;
            evecs[0] = 1.0;
            return;
        }
        davidson_solve(op, n, c0, ncv, maxiter, tol, precision, evals, evecs, stats);
    });
}

void pyci_op_destroy(pyci_op *op) { delete op; }

int pyci_add_hci(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double eps, int64_t *nadded) {
    return guarded([&] {
        PE_REQUIRE(ham->ctx == wfn->ctx, PYCI_EINVAL, "ham and wfn belong to different contexts");
        set_device(wfn->ctx);
        *nadded = add_hci_device(ham, wfn, coeffs, eps);
    });
}

int pyci_compute_enpt2(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double energy, double eps,
                       double *pt_energy, int64_t *nterms) {
    return guarded([&] {
        PE_REQUIRE(ham->ctx == wfn->ctx, PYCI_EINVAL, "ham and wfn belong to different contexts");
        PE_REQUIRE(pt_energy != nullptr, PYCI_EINVAL, "null argument");
        set_device(wfn->ctx);
        i64 nt = 0;
        const double corr = enpt2_device(ham, wfn, coeffs, energy - ham->ecore, eps, &nt);
        *pt_energy = energy + corr;
        if (nterms) *nterms = nt;
    });
}

int pyci_compute_rdms(pyci_wfn *wfn, const double *coeffs, double *rdm1, double *rdm2) {
    return guarded([&] {
        set_device(wfn->ctx);
        rdms_device(wfn, coeffs, rdm1, rdm2);
    });
}

int pyci_compute_transition_rdms(pyci_wfn *wfn1, pyci_wfn *wfn2, const double *coeffs1, const double *coeffs2,
                                 double *rdm1, double *rdm2) {
    return guarded([&] {
        PE_REQUIRE(wfn1 && wfn2 && rdm1 && rdm2, PYCI_EINVAL, "null argument");
        set_device(wfn1->ctx);
        trdms_device(wfn1, wfn2, coeffs1, coeffs2, rdm1, rdm2);
    });
}

int pyci_fanci_create(pyci_op *op, pyci_wfn *wfn, int kind, int64_t ndet_cons, const int64_t *idx_det,
                      const double *val_det, int64_t nparam_cons, const int64_t *idx_param, const double *val_param,
                      pyci_fanci **out) {
    return guarded([&] {
        PE_REQUIRE(op && wfn && out, PYCI_EINVAL, "null argument");
        PE_REQUIRE(ndet_cons >= 0 && nparam_cons >= 0, PYCI_EINVAL, "negative constraint count");
        PE_REQUIRE((ndet_cons == 0 || (idx_det && val_det)) && (nparam_cons == 0 || (idx_param && val_param)),
                   PYCI_EINVAL, "null constraint array");
        PE_REQUIRE(op->ctx == wfn->ctx, PYCI_EINVAL, "operator and wave function belong to different contexts");
        set_device(op->ctx);
        *out = fanci_create(op, wfn, kind, ndet_cons, idx_det, val_det, nparam_cons, idx_param, val_param);
    });
}

void pyci_fanci_destroy(pyci_fanci *f) {
    if (!f) return;
    (void)hipSetDevice(fanci_device(f));
    fanci_destroy(f);
}

int pyci_fanci_info(const pyci_fanci *f, int64_t *nparam, int64_t *nproj, int64_t *nconn, int64_t *nequation) {
    return guarded([&] { fanci_info(f, nparam, nproj, nconn, nequation); });
}

int pyci_fanci_overlap(pyci_fanci *f, const double *x, double *y) {
    return guarded([&] {
        PE_HIP(hipSetDevice(fanci_device(f)));
        fanci_overlap(f, x, y);
    });
}

int pyci_fanci_d_overlap(pyci_fanci *f, const double *x, double *y) {
    return guarded([&] {
        PE_HIP(hipSetDevice(fanci_device(f)));
        fanci_d_overlap(f, x, y);
    });
}

int pyci_fanci_objective(pyci_fanci *f, const double *x, double *y) {
    return guarded([&] {
        PE_HIP(hipSetDevice(fanci_device(f)));
        fanci_objective(f, x, y);
    });
}

int pyci_fanci_jacobian(pyci_fanci *f, const double *x, double *y) {
    return guarded([&] {
        PE_HIP(hipSetDevice(fanci_device(f)));
        fanci_jacobian(f, x, y);
    });
}

int pyci_batch_solve(pyci_ctx *ctx, int64_t nprob, const int64_t *nbasis, const int64_t *nocc_up,
                     const int64_t *nocc_dn, const double *ecore, const double *const *one_mo,
                     const double *const *two_mo, int64_t nroot, double tol, int precision, int64_t maxiter,
                     double *evals, double *latency_ms, int64_t *iterations, int64_t *nsigma, int *converged,
                     pyci_batch_stats *stats) {
    return guarded([&] {
        PE_REQUIRE(ctx && nbasis && nocc_up && nocc_dn && ecore && one_mo && two_mo && evals, PYCI_EINVAL,
                   "null argument");
        PE_REQUIRE(precision == PYCI_PREC_F64 || precision == PYCI_PREC_MIXED, PYCI_EINVAL, "unknown precision");
        set_device(ctx);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        batch_solve(ctx, nprob, nbasis, nocc_up, nocc_dn, ecore, one_mo, two_mo, nroot, tol, precision, maxiter,
                    evals, latency_ms, iterations, nsigma, converged, stats);
    });
}

int pyci_mlp_create(pyci_ctx *ctx, int64_t nfeat, uint64_t seed, pyci_mlp **out) {
    return guarded([&] {
        PE_REQUIRE(ctx != nullptr && out != nullptr, PYCI_EINVAL, "null argument");
        set_device(ctx);
        auto m = std::make_unique<pyci_mlp>();
        m->ctx = ctx;
        mlp_init(m.get(), (int)nfeat, seed);
        *out = m.release();
    });
}

int pyci_mlp_destroy(pyci_mlp *m) {
    return guarded([&] {
        if (!m) return;
        set_device(m->ctx);
        PE_HIP(hipStreamSynchronize(m->ctx->stream));
        delete m;
    });
}

int pyci_mlp_get_params(pyci_mlp *m, float *params) {
    return guarded([&] {
        set_device(m->ctx);
        mlp_get_params(m, params);
    });
}

int pyci_mlp_set_params(pyci_mlp *m, const float *params) {
    return guarded([&] {
        set_device(m->ctx);
        mlp_set_params(m, params);
    });
}

int pyci_mlp_predict(pyci_mlp *m, int kind, int64_t nbasis, int64_t n, const uint64_t *dets, float *y) {
    return guarded([&] {
        PE_REQUIRE(n >= 0, PYCI_EINVAL, "negative count");
        set_device(m->ctx);
        mlp_predict(m, kind, nbasis, n, dets, y);
    });
}

int pyci_mlp_loss_grad(pyci_mlp *m, int kind, int64_t nbasis, int64_t n, const uint64_t *dets, const float *target,
                       float *loss, float *grad) {
    return guarded([&] {
        set_device(m->ctx);
        *loss = mlp_loss_grad(m, kind, nbasis, n, dets, target, grad);
    });
}

int pyci_mlp_train(pyci_mlp *m, const pyci_wfn *wfn, const double *coeffs, double cmin, int64_t steps,
                   int64_t batch, double lr, uint64_t seed, float *last_loss) {
    return guarded([&] {
        PE_REQUIRE(m->ctx == wfn->ctx, PYCI_EINVAL, "network and wfn belong to different contexts");
        set_device(m->ctx);
        const float l = mlp_train(m, wfn, coeffs, cmin, steps, batch, lr, seed);
        if (last_loss) *last_loss = l;
    });
}

int pyci_add_hci_mlp(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double eps, pyci_mlp *m, double pool,
                     int64_t nadd, int score_mode, int64_t *nadded) {
    return guarded([&] {
        PE_REQUIRE(ham->ctx == wfn->ctx && m->ctx == wfn->ctx, PYCI_EINVAL,
                   "ham, wfn and network belong to different contexts");
        set_device(wfn->ctx);
        const i64 k = add_hci_mlp_device(ham, wfn, coeffs, eps, m, pool, nadd, score_mode);
        if (nadded) *nadded = k;
    });
}

}  // extern "C"
