/*
 * pyci_engine.h -- C ABI of the MI355X (gfx950) configuration-interaction engine.
 *
 * This is the drop-in boundary beneath the pybind11 module `pyci_amd._pyci`
 * (the reference's `pyci._pyci`, pyci/src/binding.cpp:29-1507). Plain pointers
 * and sizes only; no torch or pybind11 types. Every call returns 0 on success or
 * one of the PYCI_E* codes below; pyci_last_error() then holds a message for the
 * calling thread. Host pointers are never retained after a call returns; device
 * memory belongs to the handle that allocated it and is freed by its *_destroy.
 * Calls on one context are serialised on that context's HIP stream and are not
 * re-entrant; different contexts may be used from different threads.
 *
 * Determinant layout (as the reference's `dets` vectors, one 64-bit word per
 * spin string): DOCI and GenCI dets are one uint64 each; FullCI dets are two
 * uint64 words (alpha string, beta string). The GPU kernels support nbasis <= 64
 * (DOCI/FullCI spatial orbitals) and nbasis <= 64 spin-orbitals (GenCI).
 */
#ifndef PYCI_ENGINE_H
#define PYCI_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the pybind11 layer re-raises them as the reference does
 * (std::invalid_argument/domain_error -> ValueError, pybind11::type_error ->
 * TypeError, std::runtime_error -> RuntimeError; sparseop.cpp:116-136). */
#define PYCI_OK 0
#define PYCI_EINVAL 1      /* ValueError */
#define PYCI_ETYPE 2       /* TypeError */
#define PYCI_ENOCONV 3     /* RuntimeError("did not converge") */
#define PYCI_EHIP 4        /* RuntimeError: HIP runtime failure or no GPU */
#define PYCI_ENOMEM 5      /* MemoryError */
#define PYCI_EUNSUPPORTED 6 /* ValueError: shape outside the GPU kernels' limits */

/* Wave-function kinds (DOCIWfn, FullCIWfn, GenCIWfn; pyci.h:306-617). */
#define PYCI_WFN_DOCI 0
#define PYCI_WFN_FULLCI 1
#define PYCI_WFN_GENCI 2

/* Eigensolver precision modes for pyci_op_solve. */
#define PYCI_PREC_F64 0   /* FP64 sigma throughout */
#define PYCI_PREC_MIXED 1 /* FP32-value sigma, then FP64 residual-correction pass */

/* Operator storage */
#define PYCI_STORAGE_CSR 0    /* CSR in HBM (the reference's SparseOp) */
#define PYCI_STORAGE_DIRECT 1 /* matrix-free FullCI sigma; the CSR is materialised on demand */
#define PYCI_STORAGE_AUTO 2   /* DIRECT where it applies, else CSR */

typedef struct pyci_ctx pyci_ctx; /* one device + one HIP stream */
typedef struct pyci_ham pyci_ham; /* integrals resident in HBM */
typedef struct pyci_wfn pyci_wfn; /* determinant set + device index */
typedef struct pyci_op pyci_op;   /* sparse CI operator (CSR) in HBM */

/* Solver statistics filled by pyci_op_solve (may be NULL). */
typedef struct pyci_solve_stats {
    int64_t iterations;     /* Davidson iterations (all phases) */
    int64_t sigma_calls;    /* H*X products, counted per vector */
    int64_t sigma_calls_f32;/* of which FP32-value sigma */
    double residual[64];    /* final residual 2-norm per root (first 64 roots) */
} pyci_solve_stats;

/* Message of the last failing call on this thread (never NULL). */
const char *pyci_last_error(void);

/* Number of visible GPUs (0 on a host without GPUs; not an error). */
int pyci_device_count(int *count);

/* Context: binds a device and creates its stream. No counterpart in the
 * reference (its SparseOp runs on the host thread that calls it). The context
 * also keeps the eigensolver's working memory and the matrix-free operator's
 * workspace between calls (grown to the largest problem seen, released by
 * pyci_ctx_destroy), and a second stream on which the solver runs the second
 * parity sector. */
int pyci_ctx_create(int device, pyci_ctx **ctx);
void pyci_ctx_destroy(pyci_ctx *ctx);
int pyci_ctx_synchronize(pyci_ctx *ctx);
/* HIP events recorded on the context stream (for measurement). */
int pyci_ctx_timer_start(pyci_ctx *ctx);
int pyci_ctx_timer_stop(pyci_ctx *ctx, float *ms);

/* Integrals. Replaces SQuantOp(ecore, one_mo, two_mo) (squantop.cpp:163-183):
 * one_mo (n,n) and physicist two_mo (n,n,n,n) <pq|rs>, C order, float64. The
 * arrays are copied into HBM. */
int pyci_ham_create(pyci_ctx *ctx, int64_t nbasis, double ecore, const double *one_mo, const double *two_mo,
                    pyci_ham **ham);
void pyci_ham_destroy(pyci_ham *ham);

/* Determinant set. Replaces the `dets` vector and hash `dict` of OneSpinWfn /
 * TwoSpinWfn (onespinwfn.cpp:57-63,123-143; twospinwfn.cpp:57-63,129-147).
 * `dets` must hold ndet distinct determinants with the stated occupations; the
 * insertion index of each is its position. A set equal to the full space in
 * colex order is indexed by perfect colex rank, any other by a GPU hash table. */
int pyci_wfn_create(pyci_ctx *ctx, int kind, int64_t nbasis, int64_t nocc_up, int64_t nocc_dn, int64_t ndet,
                    const uint64_t *dets, pyci_wfn **wfn);
/* Appends n determinants that are not yet in the set (add_det, onespinwfn.cpp:136-143). */
int pyci_wfn_append(pyci_wfn *wfn, int64_t n, const uint64_t *dets);
int pyci_wfn_ndet(const pyci_wfn *wfn, int64_t *ndet);
/* Copies dets[start:end] to host memory (to_det_array, onespinwfn.cpp:106-110). */
int pyci_wfn_get_dets(const pyci_wfn *wfn, int64_t start, int64_t end, uint64_t *out);
/* Batched index_det (onespinwfn.cpp:123-126): -1 where absent. */
int pyci_wfn_index(const pyci_wfn *wfn, int64_t n, const uint64_t *dets, int64_t *index);
void pyci_wfn_destroy(pyci_wfn *wfn);

/* Fills the full determinant space in colex order into host memory `out`
 * (add_all_dets: onespinwfn.cpp:173-217, twospinwfn.cpp:181-245). Words per
 * spin string nword = ceil(nbasis/64) (any nbasis); FullCI writes
 * 2*nword words per det (alpha then beta), others nword. */
int pyci_fill_all_dets(pyci_ctx *ctx, int kind, int64_t nbasis, int64_t nocc_up, int64_t nocc_dn, uint64_t *out);

/* Sparse CI operator. Replaces SparseOp(ham, wfn, nrow, ncol, symmetric)
 * (sparseop.cpp:57-63 -> update :186-201 -> add_row :220-502 -> sort_row :214-218).
 * nrow/ncol < 0 mean ndet. The operator holds every structural element of rows
 * 0..nrow-1 with column < ncol, rows sorted by column; with symmetric != 0 the
 * reference-format export is the lower triangle plus diagonal. */
int pyci_op_create(pyci_ham *ham, pyci_wfn *wfn, int64_t nrow, int64_t ncol, int symmetric, pyci_op **op);
/* Same with a storage choice (PYCI_STORAGE_*). DIRECT (matrix-free sigma)
 * applies to the full FullCI space in colex order (add_all_dets), square and
 * symmetric, nbasis <= 15, both spins occupied; it stores string tables and
 * integrals instead of the CSR, which export/get_element build on first use.
 * AUTO picks DIRECT where it applies. Same operator semantics as SparseOp. */
int pyci_op_create_ex(pyci_ham *ham, pyci_wfn *wfn, int64_t nrow, int64_t ncol, int symmetric, int storage,
                      pyci_op **op);
/* Storage actually used by an operator (PYCI_STORAGE_CSR or PYCI_STORAGE_DIRECT). */
int pyci_op_storage(const pyci_op *op, int *storage);
/* update(ham, wfn) (sparseop.cpp:175-178): extends the operator to the wave
 * function's current ndet rows and columns. */
int pyci_op_update(pyci_op *op, pyci_ham *ham, pyci_wfn *wfn);
/* shape, size (reference nnz), stored nnz, symmetric flag, ecore. */
int pyci_op_info(const pyci_op *op, int64_t *nrow, int64_t *ncol, int64_t *size, int64_t *nnz_stored,
                 int *symmetric, double *ecore);
/* data()/indices()/indptr() in the reference format (sparseop.cpp:504-514):
 * size float64 values, size int64 columns, nrow+1 int64 row pointers. */
int pyci_op_export(pyci_op *op, double *data, int64_t *indices, int64_t *indptr);
/* get_element(i, j) (sparseop.cpp:89-94); 0 for j > i in symmetric mode. */
int pyci_op_get_element(pyci_op *op, int64_t i, int64_t j, double *value);
/* Diagonal H_ii for every row (host output, nrow values, ecore not added). */
int pyci_op_diagonal(pyci_op *op, double *diag);
/* y = A x for nvec vectors (perform_op / perform_op_symm, sparseop.cpp:96-112).
 * x is nvec x ncol and y nvec x nrow, row-major (one vector per row), host memory. */
int pyci_op_matvec(pyci_op *op, int64_t nvec, const double *x, double *y);
/* Same with device pointers (x, y in HBM, X row-major ncol x nvec, i.e. the nvec
 * values of one determinant are contiguous); runs on the context stream. */
int pyci_op_matvec_device(pyci_op *op, int64_t nvec, const double *x_dev, double *y_dev);
/* Allocation-free variant of pyci_op_matvec_device. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same
 * arguments and result; it calls no hipMalloc/hipFree. The matrix-free
 * operator's E workspace is a fixed per-device array of 256 MiB (held from
 * pyci_ctx_create to process exit): a call that needs more
 * (kpad * ceil(nrow * nvec / 64) * 64 * 8 bytes) fails with PYCI_EUNSUPPORTED.
 * Calls that overlap on one device share that array and race; serialise them
 * (one context stream does). The CSR product needs no workspace. */
int pyci_op_matvec_device_noalloc(pyci_op *op, int64_t nvec, const double *x_dev, double *y_dev);
/* Timed sigma for measurement: nrep launches of the SpMM kernel on device-resident
 * random X (nvec columns); returns the average kernel time from HIP events and the
 * algorithmic bytes of one launch. fp32 != 0 times the FP32-value kernel. */
int pyci_op_bench_sigma(pyci_op *op, int64_t nvec, int64_t nrep, int fp32, float *avg_ms, double *bytes);
/* Per-kernel view of the same measurement: *nparts kernels per sigma (1 for the
 * CSR; 2 for the matrix-free operator: fused gather + pair-integral product, scatter
 * pass), each with its average duration (HIP events between the kernels of each
 * of nrep sigmas), algorithmic bytes and flops per launch. Arrays hold >= 4. For the
 * matrix-free operator with nup == ndn and nvec == 1 the timed sigma is the
 * parity-sector one the eigensolver runs (product of the rows its scatter reads;
 * scatter = alpha terms + symmetrisation), on random X. */
int pyci_op_bench_sigma_parts(pyci_op *op, int64_t nvec, int64_t nrep, int fp32, int *nparts, float *avg_ms,
                              double *bytes, double *flops);
/* solve(n, c0, ncv, maxiter, tol) (sparseop.cpp:114-146,161-173): lowest n
 * eigenpairs by block Davidson on the GPU. evals (n, ecore added) and evecs
 * (n, nrow) are returned in the reference order: descending eigenvalue when
 * n > 1. c0 (nrow) or NULL. ncv/maxiter <= 0 select this library's defaults,
 * which differ from the reference's Lanczos ones (ncv = min(nrow, max(2n+1, 20)),
 * maxiter = n*nrow*10): the Davidson subspace ncv = min(nrow, max(14, 2n+1, 6n))
 * (thick restarts keep 2n Ritz vectors) and maxiter = max(1000, min(n*nrow*10,
 * 100000)) iterations. An explicit ncv is raised to at least 2n+1. */
int pyci_op_solve(pyci_op *op, int64_t n, const double *c0, int64_t ncv, int64_t maxiter, double tol, int precision,
                  double *evals, double *evecs, pyci_solve_stats *stats);
void pyci_op_destroy(pyci_op *op);

/* Heat-bath selection. Replaces add_hci (hci.cpp:252-279, per-det screening
 * :49-238): appends to `wfn` every determinant connected to some det i by an
 * excitation with |H_ji| > eps/|c_i| that is not already present. New
 * determinants are appended in first-occurrence order of the serial
 * enumeration (source det ascending, reference loop order). coeffs: ndet
 * float64 in host memory. */
int pyci_add_hci(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double eps, int64_t *nadded);

/* compute_enpt2 (enpt2.cpp:342-370, per-det generator :76-235 FullCI and
 * :257-328 GenCI; a DOCI set enters as its FullCI set of (d, d) determinants,
 * :382-386): pt_energy = energy + sum_a (sum_i H_ai c_i)^2 / (energy - ecore - H_aa)
 * over the determinants a outside the set reached from some det i by an
 * excitation with |H_ai| > eps/|c_i| (the add_hci screening). coeffs: ndet
 * float64, host memory. nterms (may be NULL): number of external determinants. */
int pyci_compute_enpt2(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double energy, double eps,
                       double *pt_energy, int64_t *nterms);

/* compute_rdms (rdm.cpp:1020-1056). coeffs: ndet float64, host memory.
 *   DOCI:   rdm1 = d0 (n,n),   rdm2 = d2 (n,n)          (rdm.cpp:20-65)
 *   FullCI: rdm1 = (2,n,n),    rdm2 = (3,n,n,n,n)      (rdm.cpp:269-530)
 *   GenCI:  rdm1 = (n,n),      rdm2 = (n,n,n,n)        (mathematically exact
 *           <a+_p a_q>, <a+_p a+_q a_s a_r>; the reference routine at
 *           rdm.cpp:532-632 has index errors, so it is not reproduced). */
int pyci_compute_rdms(pyci_wfn *wfn, const double *coeffs, double *rdm1, double *rdm2);

/* compute_transition_rdms (rdm.cpp:634-899): for every det I of wfn1 and every
 * excitation J of I present in wfn2 (and I itself), c1_I c2_J times the phase
 * into [p,q] = <1| a+_p a_q |2> and [p,q,r,s] = <1| a+_p a+_q a_s a_r |2>.
 * Shapes as pyci_compute_rdms. Both sets must have the same kind, nbasis and
 * occupations. GenCI is the exact operator algebra (the reference routine,
 * rdm.cpp:901-1009, reads coeffs2 at the wrong index and writes past its
 * arrays). coeffs1/coeffs2: host float64, one per determinant. */
int pyci_compute_transition_rdms(pyci_wfn *wfn1, pyci_wfn *wfn2, const double *coeffs1, const double *coeffs2,
                                 double *rdm1, double *rdm2);

/* ------------------------------------------------------------------------
 * Learned selector. No counterpart in the reference: it proposes which
 * heat-bath candidates add_hci (hci.cpp:252-279) appends.
 *
 * Network: occupation bits (nfeat <= 64: the determinant's bits; for FullCI
 * alpha bits 0..nbasis-1 then beta bits nbasis..2*nbasis-1) -> 256 -> 256 -> 1,
 * ReLU, predicting log10|c|. BF16 operands with FP32 accumulation on MFMA;
 * FP32 master parameters, Adam. Flat FP32 parameter layout (PYCI_MLP_NPARAM):
 *   W1[64][256] (input x hidden; rows >= nfeat are zero), b1[256],
 *   W2[256][256], b2[256], w3[256], b3.
 * Forward: h1 = bf16(relu(bf16(x) W1b + b1)), h2 = bf16(relu(h1 W2b + b2)),
 * y = h2 . w3 + b3, with W1b, W2b the bf16-rounded weights.
 * ------------------------------------------------------------------------ */
#define PYCI_MLP_HIDDEN 256
#define PYCI_MLP_NPARAM (64 * 256 + 256 + 256 * 256 + 256 + 256 + 1)
typedef struct pyci_mlp pyci_mlp;

/* Creates a network with seeded He-uniform weights (zero biases). */
int pyci_mlp_create(pyci_ctx *ctx, int64_t nfeat, uint64_t seed, pyci_mlp **out);
int pyci_mlp_destroy(pyci_mlp *m);
int pyci_mlp_get_params(pyci_mlp *m, float *params);       /* PYCI_MLP_NPARAM floats */
int pyci_mlp_set_params(pyci_mlp *m, const float *params); /* resets the Adam state */
/* y[n] = network(dets) for host determinants (nkey words each: 2 for FullCI). */
int pyci_mlp_predict(pyci_mlp *m, int kind, int64_t nbasis, int64_t n, const uint64_t *dets, float *y);
/* Mean-squared-error loss over n samples and its gradient (PYCI_MLP_NPARAM floats). */
int pyci_mlp_loss_grad(pyci_mlp *m, int kind, int64_t nbasis, int64_t n, const uint64_t *dets,
                       const float *target, float *loss, float *grad);
/* Trains on the wave function's determinants with target log10|coeffs| (host
 * coeffs, ndet float64; determinants with |c| < cmin are left out): `steps`
 * Adam updates on minibatches of `batch` samples drawn with a counter-based
 * generator from `seed`. Returns the loss of the last minibatch. */
int pyci_mlp_train(pyci_mlp *m, const pyci_wfn *wfn, const double *coeffs, double cmin, int64_t steps,
                   int64_t batch, double lr, uint64_t seed, float *last_loss);

/* add_hci with the learned selector: the heat-bath pool is taken at eps/pool
 * (pool >= 1); `nadd` < 0 appends as many determinants as plain add_hci(eps)
 * would, otherwise min(nadd, pool size); candidates are ranked by the network
 * (score_mode 0) or by network + log10 of the heat-bath score (1); ties keep
 * the first-occurrence order. nadded may be NULL. */
int pyci_add_hci_mlp(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs, double eps, pyci_mlp *m, double pool,
                     int64_t nadd, int score_mode, int64_t *nadded);

/* ------------------------------------------------------------------------
 * FanCI geminal objectives. Replace AP1roGObjective / APIGObjective
 * (fanci.cpp:21-217, ap1rog.cpp:21-220, apig.cpp:21-195): the projected
 * Schrodinger equations f_n = <n|H|Psi> - E <n|Psi> over the P space, their
 * Jacobian, and the geminal overlaps (permanents) over the S space.
 * ------------------------------------------------------------------------ */
#define PYCI_FANCI_AP1ROG 0 /* parameters c[i][a]: nocc x (nbasis - nocc), reference = first nocc orbitals */
#define PYCI_FANCI_APIG 1   /* parameters c[p][i]: nbasis x nocc */
typedef struct pyci_fanci pyci_fanci;

/* op: a non-symmetric CSR operator with nproj rows and nconn columns over the
 * DOCI wave function wfn (its first nconn determinants are the S space, the
 * first nproj the P space); the operator must outlive the objective. Optional
 * constraints (counts may be 0, pointers then NULL): determinant overlaps
 * c[idx_det[i]] - val_det[i] and parameters x[idx_param[i]] - val_param[i]. */
int pyci_fanci_create(pyci_op *op, pyci_wfn *wfn, int kind, int64_t ndet_cons, const int64_t *idx_det,
                      const double *val_det, int64_t nparam_cons, const int64_t *idx_param, const double *val_param,
                      pyci_fanci **out);
void pyci_fanci_destroy(pyci_fanci *f);
/* nparam excludes the energy; nequation = nproj + constraints. */
int pyci_fanci_info(const pyci_fanci *f, int64_t *nparam, int64_t *nproj, int64_t *nconn, int64_t *nequation);
/* overlap (Objective::py_overlap, fanci.cpp:226-231): y[nconn] from x[nparam]. */
int pyci_fanci_overlap(pyci_fanci *f, const double *x, double *y);
/* d_overlap (fanci.cpp:233-239): y column-major nconn x nparam. */
int pyci_fanci_d_overlap(pyci_fanci *f, const double *x, double *y);
/* objective (fanci.cpp:129-160): x[nparam + 1] (energy last) -> y[nequation]. */
int pyci_fanci_objective(pyci_fanci *f, const double *x, double *y);
/* jacobian (fanci.cpp:162-217): y column-major nequation x (nparam + 1); the
 * parameter columns come from one product of the operator with all derivative
 * columns instead of one perform_op per parameter (fanci.cpp:197-204). */
int pyci_fanci_jacobian(pyci_fanci *f, const double *x, double *y);

/* ------------------------------------------------------------------------
 * Batched solves of many uneven FullCI problems. No counterpart in the
 * reference: each problem is SparseOp(ham, wfn) + solve_ci(n) (sparseop.cpp:
 * 114-146) on the full FullCI space of its own integrals (add_all_dets). The
 * problems below 150,000 determinants run together on the device (matrix-free
 * sigma, one kernel launch per Davidson step over all of them, the iteration
 * replayed as a hipGraph); the larger ones are solved by the single-problem
 * solver beside them, two at a time from the largest down (the second on a
 * context the call creates and destroys).
 * ------------------------------------------------------------------------ */
typedef struct pyci_batch_stats {
    int64_t graph_launches; /* Davidson iterations replayed (the graph's maximum) */
    int64_t rows_per_block; /* autotuned rows per workgroup of the tall-skinny kernels */
    int64_t converged;      /* problems that converged */
    double setup_ms;        /* host set-up of the graph path (tables, integrals, diagonals) */
    double total_ms;        /* whole call */
    int64_t single_path;    /* problems solved by the single-problem solver */
} pyci_batch_stats;

/* nprob problems: nbasis[p] <= 15 spatial orbitals, nocc_up[p] >= nocc_dn[p] >= 1,
 * ecore[p], one_mo[p] (n,n) and two_mo[p] (n,n,n,n) host arrays as for
 * pyci_ham_create. nroot <= 8 lowest roots each, tol as pyci_op_solve, precision
 * PYCI_PREC_F64 or PYCI_PREC_MIXED. Outputs (host, may be NULL except evals):
 * evals[p * nroot + j] (ecore added; descending for nroot > 1 as solve_ci),
 * latency_ms[p] from the first iteration of the problem's path to its
 * convergence (device clock), iterations[p], nsigma[p] (vectors multiplied by H),
 * converged[p] (an unconverged problem of the single-problem path reports NaN
 * roots). */
int pyci_batch_solve(pyci_ctx *ctx, int64_t nprob, const int64_t *nbasis, const int64_t *nocc_up,
                     const int64_t *nocc_dn, const double *ecore, const double *const *one_mo,
                     const double *const *two_mo, int64_t nroot, double tol, int precision, int64_t maxiter,
                     double *evals, double *latency_ms, int64_t *iterations, int64_t *nsigma, int *converged,
                     pyci_batch_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* PYCI_ENGINE_H */
