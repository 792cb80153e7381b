// Internal definitions shared by the engine translation units (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pyci_engine.h"

namespace pe {

using u64 = uint64_t;
using i64 = int64_t;

// ---------------------------------------------------------------------------
// Errors: host code throws EngineError; the C ABI converts it into a status.
// ---------------------------------------------------------------------------
struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string &msg);

#define PE_HIP(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw ::pe::EngineError(PYCI_EHIP, std::string(#x " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define PE_STR2(x) #x
#define PE_STR(x) PE_STR2(x)
#define PE_LAUNCH_CHECK()                                                                                \
    do {                                                                                                 \
        hipError_t e_ = hipGetLastError();                                                               \
        if (e_ != hipSuccess)                                                                            \
            throw ::pe::EngineError(PYCI_EHIP, std::string("kernel launch at " __FILE__ ":" PE_STR(__LINE__) \
                                                           " failed: ") + hipGetErrorString(e_));         \
    } while (0)

#define PE_REQUIRE(cond, code, msg)                     \
    do {                                                \
        if (!(cond)) throw ::pe::EngineError(code, msg); \
    } while (0)

// Device buffer owned by RAII.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p, n = o.n;
            o.p = nullptr, o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    // at least count elements: keeps the allocation when it is large enough
    void ensure(size_t count) {
        if (!p || n < count) alloc(count);
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw EngineError(e == hipErrorOutOfMemory ? PYCI_ENOMEM : PYCI_EHIP,
                              "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                  hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) {
            // a failed free must not linger as the thread's last HIP error, where it
            // would be reported against an unrelated later launch
            if (hipFree(p) != hipSuccess) (void)hipGetLastError();
        }
        p = nullptr;
        n = 0;
    }
};

// Page-locked host buffer (hipHostMalloc): copies to and from it run
// asynchronously on the stream instead of staging through a pageable bounce.
template <typename T>
struct HostBuf {
    T *p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { release(); }
    void ensure(size_t count) {
        if (!p || n < count) alloc(count);
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            throw EngineError(e == hipErrorOutOfMemory ? PYCI_ENOMEM : PYCI_EHIP,
                              "hipHostMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                  hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p && hipHostFree(p) != hipSuccess) (void)hipGetLastError();
        p = nullptr;
        n = 0;
    }
};

}  // namespace pe

namespace pe {
// Working memory of one Davidson run (davidson.hip). The context keeps one set per
// parity-sector stream between solves, so that a solve allocates nothing once the
// sizes have been seen; the sets are released with the context.
struct DavBufs {
    DevBuf<double> V, W, X, WX, T, rowbuf, rowout, partial, small, RV, RW, dv, pb, dsym;
    DevBuf<float> row32, out32;
    HostBuf<double> hb, hdiag;
};
}  // namespace pe

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------
struct pyci_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    pe::DevBuf<pe::i64> binom;  // C(n, k) for 0 <= n, k <= 64, row-major [n][k] (65 x 65)
    pe::DevBuf<unsigned char> mf_ws;  // E workspace of the matrix-free sigma (shared by the operators)
    // second stream and workspace: the solver runs the second parity sector on them,
    // concurrently with the first (created on first use)
    hipStream_t stream2 = nullptr;
    pe::DevBuf<unsigned char> mf_ws2;
    std::unique_ptr<pe::DavBufs> dav[2];  // solver working memory per sector stream
    double *noalloc_ws = nullptr;     // the fixed workspace of the allocation-free product on this device
};

struct pyci_ham {
    pyci_ctx *ctx = nullptr;
    pe::i64 nbasis = 0;
    double ecore = 0.0;
    pe::DevBuf<double> one_mo, two_mo;
};

struct pyci_wfn {
    pyci_ctx *ctx = nullptr;
    int kind = 0;
    pe::i64 nbasis = 0, nup = 0, ndn = 0;
    int nkey = 1;  // words per determinant
    pe::i64 ndet = 0;
    pe::DevBuf<pe::u64> dets;  // capacity >= ndet * nkey
    // index: perfect colex rank when the set is the full space in colex order
    bool colex = false;
    pe::i64 nsd = 1;  // beta (or only) string count, for FullCI colex rank
    // otherwise: open-addressing hash (key words per slot = 2, value = index, -1 empty)
    // with one tag byte per slot (0 empty, else 0x80 | the hash's top 7 bits), 8 per word
    pe::DevBuf<pe::u64> hkeys;
    pe::DevBuf<pe::i64> hvals;
    pe::DevBuf<pe::u64> htags;
    pe::i64 hcap = 0;
    // large tables only: a presence filter, one bit per determinant's filter hash in an
    // array of fbits bits (about 4 per determinant), for lookups that miss
    pe::DevBuf<unsigned> hfilt;
    pe::i64 fbits = 0;
};

namespace pe {
// Matrix-free FullCI operator state (direct.hip).
struct DirectOp {
    int n = 0, nup = 0, ndn = 0, npair = 0, kpad = 0, nea = 0, neb = 0;
    i64 na = 0, nbs = 0;         // alpha / beta string counts
    double ecore = 0.0;
    bool csr_ready = false;      // the CSR arrays of the operator have been materialised
    DevBuf<int2> tab_a, tab_b;   // [string][entry] {rank of the excited string, pair | negative << 16}
    DevBuf<int2> tab;            // tab_a rows, tab_b entries slot-major (entry t of every beta string together), then the sentinel
    DevBuf<u64> str_a, str_b;    // strings in colex order
    // parity-sector layout of E (nup == ndn, where it saves work): only the pairs the alpha
    // scatter reads, per alpha string; ml = their row tiles (0: layout not used)
    int ml = 0;
    i64 nbp = 0;                 // beta strings rounded up to 64
    DevBuf<int> live;            // [alpha string][16 ml] pair of each live row (pad: a zero row of V)
    DevBuf<int2> tab_al;         // tab_a with {rank of the excited string, live row | negative << 16}
    DevBuf<double> one_mo, two_mo, V64, k64;
    DevBuf<float> V32, k32;
};
}  // namespace pe

// Learned selector (mlp.hip): FP32 master parameters + Adam state, bf16
// operand copies of the weights, and workspaces sized for the largest batch.
struct pyci_mlp {
    pyci_ctx *ctx = nullptr;
    int nfeat = 0;
    long long adam_t = 0;
    pe::DevBuf<float> params, grads, adam_m, adam_v;
    pe::DevBuf<uint16_t> w1t;  // [256][64]  layer-1 operand: row = hidden unit, k = input feature
    pe::DevBuf<uint16_t> w2t;  // [256][256] layer-2 operand: row = output unit f2, k = f1
    pe::DevBuf<uint16_t> w2;   // [256][256] data-backward operand: row = f1, k = f2
    // workspaces (grown on demand)
    pe::i64 cap = 0;           // rows the workspaces hold (multiple of 128)
    pe::DevBuf<uint16_t> x, xt, h1, h1t, h2t, dz2, dz2t, dz1t;
    pe::DevBuf<float> y, ypart, dy, tgt, wpart, red;
};

struct pyci_op {
    pyci_ctx *ctx = nullptr;
    int kind = 0;
    pe::i64 nrow = 0, ncol = 0;
    int symmetric = 1;
    double ecore = 0.0;
    pe::i64 nnz = 0;       // stored (full-pattern) nonzeros
    pe::i64 size_ref = 0;  // reference-format nonzeros (lower+diag if symmetric)
    pe::DevBuf<pe::i64> rowptr;   // nrow+1
    pe::DevBuf<pe::i64> lowcnt;   // per-row count of entries with col <= row (symmetric export)
    pe::DevBuf<int32_t> cols;     // nnz
    pe::DevBuf<double> vals;      // nnz
    pe::DevBuf<float> vals32;     // nnz, built on demand for the FP32 sigma
    pe::DevBuf<double> diag;      // nrow
    // rows built in segments for non-symmetric update(): [row0, row1) with its ncol
    std::vector<std::pair<pe::i64, pe::i64>> segments;
    // storage PYCI_STORAGE_DIRECT: matrix-free sigma; the CSR arrays above are
    // filled only on demand (exports, element queries)
    int storage = PYCI_STORAGE_CSR;
    std::unique_ptr<pe::DirectOp> direct;
    // beta-string count when the rows are the full FullCI space in colex order with
    // nup == ndn (the alpha <-> beta exchange is then J = a nb + b -> b nb + a), else 0
    pe::i64 swap_nb = 0;
};

namespace pe {

// ---------------------------------------------------------------------------
// Device views
// ---------------------------------------------------------------------------
struct WfnView {
    int kind, nkey;
    int nb, nup, ndn;  // nup = nocc for one-spin kinds
    i64 ndet;
    const u64 *dets;
    int colex;
    i64 nsd;
    const i64 *binom;  // 65 x 65
    const u64 *hkeys;
    const i64 *hvals;
    const u64 *htags;
    const unsigned *hfilt;  // or null
    u64 hmask, fmask;
    // GenCI row builds: every determinant of the set has the same number of occupied
    // orbitals below szh (0: not known to), so an excitation that changes that number
    // cannot find a column and is not enumerated
    int szh = 0;
};

struct HamView {
    int n;
    const double *one, *two;  // one (n,n), two (n,n,n,n)
};

WfnView view_of(const pyci_wfn *w);
HamView view_of(const pyci_ham *h);

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__host__ __device__ inline u64 mix64(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// Slot hash of determinant (a, b) in the device index: one finaliser over a and an
// odd multiple of b (b = 0 for one-spin kinds, where the multiply folds away: the
// GenCI row builder probes ~42,000 candidates per row, and the hash is a large part
// of each probe's instructions).
__host__ __device__ inline u64 hash_key(u64 a, u64 b) {
    return mix64((a + 0x9e3779b97f4a7c15ULL) ^ (b * 0xd6e8feb86659fd93ULL));
}

__device__ inline i64 colex_rank(u64 det, const i64 *binom) {
    i64 r = 0;
    int k = 0;
    while (det) {
        int p = __builtin_ctzll(det);
        ++k;
        r += binom[p * 65 + k];
        det &= det - 1;
    }
    return r;
}

// Index of determinant (a, b) (b = 0 for one-spin kinds), or -1.
__device__ inline i64 wfn_lookup(const WfnView &w, u64 a, u64 b) {
    if (w.colex) {
        if (w.nkey == 2) return colex_rank(a, w.binom) * w.nsd + colex_rank(b, w.binom);
        return colex_rank(a, w.binom);
    }
    // Linear probing over the tag bytes, eight slots per load: a probe run ends at the
    // first empty byte, and only slots whose tag matches are compared by key, so a
    // determinant that is not in the set usually costs one 8-byte load. Byte tests are
    // word-wide: for x with bytes 0 or >= 0x80, (x - 0x0101..) & ~x & 0x8080.. flags
    // exactly the zero bytes; applied to x ^ tag it flags the matching bytes (and at
    // most spurious bytes above a true match, which the key compare rejects).
    const u64 hk = hash_key(a, b), ones = 0x0101010101010101ULL, highs = 0x8080808080808080ULL;
    if (w.hfilt) {
        // bits 24..55 of the hash: independent of the slot (low bits) and the tag (top 7)
        const u64 f = (hk >> 24) & w.fmask;
        if (!((w.hfilt[f >> 5] >> (f & 31)) & 1u)) return -1;
    }
    const u64 tagw = ((hk >> 57) | 0x80ULL) * ones;
    u64 h = hk & w.hmask;
    for (;;) {
        const u64 x = w.htags[h >> 3];
        const int j0 = (int)(h & 7);
        const u64 keep = ~0ULL << (8 * j0);  // slots h .. end of the word
        const u64 empty = ((x - ones) & ~x & highs) & keep;
        const u64 y = x ^ tagw;
        u64 match = ((y - ones) & ~y & highs) & keep;
        // matches before the first empty slot, in probe order
        if (empty) match &= (empty & (~empty + 1)) - 1;
        while (match) {
            const u64 s = (h & ~7ULL) + (u64)(__builtin_ctzll(match) >> 3);
            if (w.hkeys[2 * s] == a && w.hkeys[2 * s + 1] == b) return w.hvals[s];
            match &= match - 1;
        }
        if (empty) return -1;
        h = ((h | 7ULL) + 1ULL) & w.hmask;
    }
}

// Number of set bits of det strictly between lo and hi (lo < hi < 64).
__device__ inline int bits_between(u64 det, int lo, int hi) {
    u64 below = (hi >= 64) ? ~0ULL : ((1ULL << hi) - 1ULL);
    u64 above = (lo >= 63) ? 0ULL : ~((1ULL << (lo + 1)) - 1ULL);
    return __popcll(det & below & above);
}

// phase_single_det (common.cpp:163-194).
__device__ inline int phase1(u64 det, int i, int a) {
    return (bits_between(det, min(i, a), max(i, a)) & 1) ? -1 : 1;
}

// phase_double_det (common.cpp:196-261), counts on the original determinant.
__device__ inline int phase2(u64 det, int i1, int i2, int a1, int a2) {
    int n = bits_between(det, min(i1, a1), max(i1, a1)) + bits_between(det, min(i2, a2), max(i2, a2));
    if (i2 < a1 || i1 > a2) ++n;
    return (n & 1) ? -1 : 1;
}

__device__ inline u64 full_mask(int nb) { return nb >= 64 ? ~0ULL : ((1ULL << nb) - 1ULL); }

// Pair index t = k (k - 1) / 2 + i with i < k, decoded.
__device__ inline void tri_decode(int t, int &i, int &k) {
    int kk = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
    while (kk * (kk - 1) / 2 > t) --kk;
    while ((kk + 1) * kk / 2 <= t) ++kk;
    k = kk;
    i = t - kk * (kk - 1) / 2;
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
i64 binom_host(i64 n, i64 k);
inline unsigned grid_for(i64 n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

// Exclusive scan of n int64 counts in place on the stream; returns the total.
i64 exclusive_scan(pyci_ctx *ctx, i64 *d_data, i64 n);

// CSR build entry points (csr_build.hip)
void build_rows(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn, i64 row0, i64 row1, i64 ncol,
                DevBuf<i64> &rowptr_seg, DevBuf<int32_t> &cols, DevBuf<double> &vals, DevBuf<i64> &lowcnt_seg,
                DevBuf<double> &diag_seg, i64 &nnz_seg, i64 &low_total);
// SpMM (sigma.hip): Y = A X, X ncol x nvec row-major, Y nrow x nvec row-major
// sector = +1 / -1: the vectors lie in that parity sector of the alpha <-> beta
// exchange (a hint the matrix-free operator uses; 0 = any vectors)
// ws (optional): the matrix-free operator's E workspace (default: the context's mf_ws)
void sigma_f64(const pyci_op *op, int nvec, const double *X, double *Y, hipStream_t s, int sector = 0,
               DevBuf<unsigned char> *ws = nullptr);
void sigma_f32(const pyci_op *op, int nvec, const float *X, float *Y, hipStream_t s, int sector = 0,
               DevBuf<unsigned char> *ws = nullptr);
void ensure_vals32(pyci_op *op);
// Davidson (davidson.hip)
void davidson_solve(pyci_op *op, i64 n, const double *c0, i64 ncv, i64 maxiter, double tol, int precision,
                    double *evals, double *evecs, pyci_solve_stats *stats);
// index (index.hip)
void wfn_build_index(pyci_wfn *w);
void fill_all_dets_device(pyci_ctx *ctx, int kind, i64 nb, i64 nup, i64 ndn, u64 *d_out, int nword);
// HCI / RDMs
i64 add_hci_device(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs_host, double eps);
void rdms_device(pyci_wfn *wfn, const double *coeffs_host, double *rdm1, double *rdm2);
void trdms_device(pyci_wfn *w1, pyci_wfn *w2, const double *c1_host, const double *c2_host, double *rdm1,
                  double *rdm2);
// ENPT2 correction sum_a (sum_i H_ai c_i)^2 / (e - H_aa), e = energy - ecore (hci.hip)
double enpt2_device(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs_host, double e, double eps, i64 *nterms);

// Heat-bath candidate pool (hci.hip): the new determinants passing eps_pool in
// first-occurrence order, their heat-bath scores max_i |H_ji c_i|, and whether
// plain add_hci(eps) would take each (nplain of them). dc: device coefficients.
struct HciPool {
    DevBuf<u64> dets;
    DevBuf<double> hb;
    DevBuf<int> plain;
    i64 n = 0, nplain = 0;
};
void hci_pool_device(pyci_ham *ham, pyci_wfn *wfn, const double *dc, double eps, double eps_pool, HciPool &P);

// Matrix-free FullCI operator (direct.hip)
// What the sigma passes read of a matrix-free operator: the alpha and beta string
// tables, the merged table with its sentinel, counts.
struct MfView {
    const int2 *tab, *tab_a, *tab_b;
    i64 na, nbs;
    int npair, kpad, nea, neb;
    // parity-sector live-row layout (ml 0: none)
    int ml = 0;
    i64 nbp = 0;
    const int *live = nullptr;
    const int2 *tab_al = nullptr;
};
template <typename T>
void mf_sigma_batched(const MfView &M, i64 nd, int nvec_max, const T *X, T *Y, T *W, const T *V, const T *kv,
                      const long long *dyn, int sector, hipStream_t s);
// pair dimension kpad -> k extent of the fused pass (its sink row index)
int mf_sink_row(int kpad);
bool direct_applicable(const pyci_ham *ham, const pyci_wfn *wfn, i64 nrow, i64 ncol, int symmetric);
void direct_setup(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn);
template <typename T>
void direct_sigma(const pyci_op *op, int nvec, const T *X, T *Y, hipStream_t s, hipEvent_t *ev = nullptr,
                  int sector = 0, DevBuf<unsigned char> *ws = nullptr);
void direct_bench_parts(pyci_op *op, i64 nvec, i64 nrep, int fp32, float *ms, double *bytes, double *flops);
void direct_ensure_csr(pyci_op *op);
// Allocation-free matrix-free product (direct.hip): E in a fixed per-device array
double *noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
void direct_sigma_noalloc(const pyci_op *op, int nvec, const double *X, double *Y, double *ws, hipStream_t s);
// helpers shared with the batched solver (direct.hip)
std::vector<u64> colex_strings(int n, int k);
i64 rank_host(u64 det, const std::vector<i64> &binom);
std::vector<int2> string_table(int n, const std::vector<u64> &S, int ne, const std::vector<i64> &binom);
void pair_integrals(int n, const double *one, const double *two, int kp, std::vector<double> &V,
                    std::vector<double> &kv);
void mf_diag_launch(i64 nd, i64 NB, const u64 *sa, const u64 *sb, HamView H, double *diag, hipStream_t s);
void mf_wallclock_launch(unsigned long long *t, hipStream_t s);
// Batched solver (batch.hip)
void batch_solve(pyci_ctx *ctx, i64 nprob, const i64 *nbasis, const i64 *nup, const i64 *ndn, const double *ecore,
                 const double *const *one_mo, const double *const *two_mo, i64 nroot, double tol, int precision,
                 i64 maxiter, double *evals, double *latency_ms, i64 *iters, i64 *nsigma, int *converged,
                 pyci_batch_stats *stats);

// FanCI geminal objectives (fanci.hip)
pyci_fanci *fanci_create(pyci_op *op, pyci_wfn *wfn, int kind, i64 ndc, const i64 *idc, const double *vdc, i64 npc,
                         const i64 *ipc, const double *vpc);
void fanci_info(const pyci_fanci *F, i64 *nparam, i64 *nproj, i64 *nconn, i64 *nequation);
int fanci_device(const pyci_fanci *F);
void fanci_destroy(pyci_fanci *F);
void fanci_overlap(pyci_fanci *F, const double *x, double *y);
void fanci_d_overlap(pyci_fanci *F, const double *x, double *y);
void fanci_objective(pyci_fanci *F, const double *x, double *y);
void fanci_jacobian(pyci_fanci *F, const double *x, double *y);

// Learned selector (mlp.hip)
void mlp_init(pyci_mlp *m, int nfeat, u64 seed);
void mlp_set_params(pyci_mlp *m, const float *host);
void mlp_get_params(pyci_mlp *m, float *host);
void mlp_predict(pyci_mlp *m, int kind, i64 nbasis, i64 n, const u64 *dets, float *y);
void mlp_predict_device(pyci_mlp *m, int kind, i64 nbasis, const u64 *d_dets, i64 n, float *d_y);
float mlp_loss_grad(pyci_mlp *m, int kind, i64 nbasis, i64 n, const u64 *dets, const float *target, float *grad);
float mlp_train(pyci_mlp *m, const pyci_wfn *wfn, const double *coeffs, double cmin, i64 steps, i64 batch, double lr,
                u64 seed);
i64 add_hci_mlp_device(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs_host, double eps, pyci_mlp *m,
                       double pool, i64 nadd, int mode);

}  // namespace pe
