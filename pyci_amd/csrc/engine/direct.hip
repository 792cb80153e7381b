// Matrix-free FullCI operator over the full determinant space in colex order.
//
// The reference always stores the CSR (SparseOp::update, sparseop.cpp:186-201)
// and multiplies with it (perform_op_symm, sparseop.cpp:148-159). For the full
// FullCI space this operator applies the same Hamiltonian without storing it,
// in the spin-free form
//   H = sum_pq k_pq E_pq + 1/2 sum_pqrs (pq|rs) E_pq E_rs,
//   k_pq = h_pq - 1/2 sum_r (pr|rq),  E_pq = E^alpha_pq + E^beta_pq,
// with (pq|rs) = two_mo[p][r][q][s]. Determinant J = (Ja, Jb) has index
// Ja * NB + Jb (colex ranks of its strings), so a vector is a row-major
// NA x NB matrix. One sigma Y = H X is three passes:
//   D[kl][J] = sum_K <J|E_kl|K> X[K]                 (gathers over single-excitation tables)
//   E[ij][J] = 1/2 sum_kl (ij|kl) D[kl][J] + k_ij X[J] (dense product on MFMA, in place)
//   Y[I]     = sum_ij sum_J <I|E_ij|J> E[ij][J]        (gathers over the same tables)
// Both D and E are symmetric in their orbital pair ((ij|kl) = (ji|kl) for real
// orbitals), so pairs are packed as p = i (i + 1) / 2 + j, i >= j, and D holds
// D_kl + D_lk. Each string's table lists, for k occupied and l virtual or l = k,
// the rank of the string with k replaced by l, the pair of (k, l) and the sign
// (-1)^(occupied orbitals strictly between k and l) (phase_single_det,
// common.cpp:163-194); entries are sorted by pair, so that the D pass writes
// each D element exactly once by merging the alpha and beta lists.
#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "internal.hpp"

namespace pe {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Determinant of this thread. TILED: the 256 threads of a block cover a 16 x 16
// tile (alpha rank x beta rank) of the string matrix, and tiles of one column range
// run on one XCD (blocks go round-robin over the 8 XCDs): the alpha gathers, which
// walk columns of x, then find them in that XCD's L2. Otherwise 256 consecutive
// determinants per block.
template <bool TILED>
__device__ inline bool det_of(i64 nd, i64 NA, i64 NB, i64 &Ja, i64 &Jb) {
    if constexpr (TILED) {
        const i64 nta = (NA + 15) / 16, ntb = (NB + 15) / 16, slots = (ntb + 7) / 8;
        const i64 b = blockIdx.x, xcd = b % 8, i = b / 8;
        const i64 tcol = xcd + 8 * (i % slots), trow = i / slots;
        if (tcol >= ntb || trow >= nta) return false;
        Ja = trow * 16 + (threadIdx.x >> 4);
        Jb = tcol * 16 + (threadIdx.x & 15);
        return Ja < NA && Jb < NB;
    } else {
        const i64 J = (i64)blockIdx.x * blockDim.x + threadIdx.x;
        if (J >= nd) return false;
        Ja = J / NB;
        Jb = J - Ja * NB;
        return true;
    }
}

static inline unsigned tiled_grid(i64 NA, i64 NB) {
    const i64 nta = (NA + 15) / 16, ntb = (NB + 15) / 16;
    return (unsigned)(8 * ((ntb + 7) / 8) * nta);
}

// D pass: one thread per determinant J; NV >= nvec vectors (row-major X[J][v]).
template <int NV, typename T, bool TILED>
__global__ void __launch_bounds__(256) k_mf_d(i64 nd, i64 NB, int npair, int nea, int neb, const int2 *__restrict__ ta,
                                              const int2 *__restrict__ tb, int nvec, const T *__restrict__ X,
                                              T *__restrict__ W, i64 ldw) {
    i64 Ja, Jb;
    if (!det_of<TILED>(nd, nd / NB, NB, Ja, Jb)) return;
    const i64 J = Ja * NB + Jb;
    const int2 *pa = ta + Ja * nea, *pb = tb + Jb * neb;
    int ia = 0, ib = 0;
    int2 ea = pa[0], eb = pb[0];
    for (int q = 0; q < npair; ++q) {
        T acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = (T)0;
        if (ia < nea && (ea.y & 0xffff) == q) {
            const T s = (ea.y >> 16) ? (T)-1 : (T)1;
            const T *x = X + ((i64)ea.x * NB + Jb) * nvec;
#pragma unroll
            for (int v = 0; v < NV; ++v)
                if (v < nvec) acc[v] += s * x[v];
            if (++ia < nea) ea = pa[ia];
        }
        if (ib < neb && (eb.y & 0xffff) == q) {
            const T s = (eb.y >> 16) ? (T)-1 : (T)1;
            const T *x = X + (Ja * NB + eb.x) * nvec;
#pragma unroll
            for (int v = 0; v < NV; ++v)
                if (v < nvec) acc[v] += s * x[v];
            if (++ib < neb) eb = pb[ib];
        }
        T *w = W + (i64)q * ldw + J * nvec;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            if (v < nvec) w[v] = acc[v];
    }
}

// Product pass, in place on W (kpad rows, ldw columns, ldw % 128 == 0):
//   W[p][c] = 1/2 sum_q V[p][q] W[q][c] + k[p] X[c]      (X[c] = 0 for c >= ncol)
// V is kpad x kpad (zero padded), MT = kpad / 16 row tiles. 256 threads: each
// wave owns 32 columns (two 16-column tiles) and all MT row tiles; every column
// is read completely before the wave writes it. MFMA 16x16x4: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator r of lane l is
// D[4 (l / 16) + r][l % 16] for the FP32 instruction and D[l / 16 + 4 r][l % 16]
// for the FP64 one (lane layouts checked on the device).
template <int MT, typename T>
__global__ void __launch_bounds__(256) k_mf_gemm(i64 ncol, const T *__restrict__ V, const T *__restrict__ kv,
                                                 const T *__restrict__ X, T *W, i64 ldw) {
    constexpr int KP = MT * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const i64 c0 = (i64)blockIdx.x * 128 + wave * 32;
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
    acc_t acc[MT][2];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        acc[m][0] = acc_t{};
        acc[m][1] = acc_t{};
    }
    // columns at or beyond ncol are read as zero: they stay zero in every row, so
    // the zero pad rows never meet a non-finite value
    const bool live0 = c0 + lr < ncol, live1 = c0 + 16 + lr < ncol;
#pragma unroll 2
    for (int k0 = 0; k0 < KP; k0 += 4) {
        const T b0 = live0 ? W[(i64)(k0 + lk) * ldw + c0 + lr] : (T)0;
        const T b1 = live1 ? W[(i64)(k0 + lk) * ldw + c0 + 16 + lr] : (T)0;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const T a = V[(16 * m + lr) * KP + k0 + lk];
            if constexpr (sizeof(T) == 8) {
                acc[m][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[m][0], 0, 0, 0);
                acc[m][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[m][1], 0, 0, 0);
            } else {
                acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0, acc[m][0], 0, 0, 0);
                acc[m][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1, acc[m][1], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const i64 c = c0 + 16 * t + lr;
        const T x = (c < ncol) ? X[c] : (T)0;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = (sizeof(T) == 8) ? 16 * m + lk + 4 * r : 16 * m + 4 * lk + r;
                W[(i64)p * ldw + c] = (T)0.5 * acc[m][t][r] + kv[p] * x;
            }
    }
}

// Product pass, second form: each wave owns 16 columns and first loads all of its
// B fragments (KP / 4 per lane) so that every load is in flight at once; V (and k)
// sit in LDS when VLDS (kpad^2 values, row stride kpad + 1 against bank conflicts),
// else they are read through the caches. 256 threads = 64 columns per block.
template <int MT, typename T, bool VLDS>
__global__ void __launch_bounds__(256) k_mf_gemm2(i64 ncol, const T *__restrict__ V, const T *__restrict__ kv,
                                                  const T *__restrict__ X, T *W, i64 ldw) {
    constexpr int KP = MT * 16, NS = KP / 4, LS = KP + 1;
    extern __shared__ unsigned char smem_raw[];
    T *vs = reinterpret_cast<T *>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    if constexpr (VLDS) {
        // V once per block; the block then walks column tiles of 64 (persistent grid)
        for (int i = threadIdx.x; i < KP * KP; i += 256) vs[(i / KP) * LS + (i % KP)] = V[i];
        __syncthreads();
    }
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
    const i64 ntiles = (ncol + 63) / 64;
    for (i64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const i64 c = tile * 64 + wave * 16 + lr;
        const bool live = c < ncol;
        T b[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) b[q] = live ? W[(i64)(4 * q + lk) * ldw + c] : (T)0;
        acc_t acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = acc_t{};
#pragma unroll
        for (int q = 0; q < NS; ++q) {
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const T a = VLDS ? vs[(16 * m + lr) * LS + 4 * q + lk] : V[(16 * m + lr) * KP + 4 * q + lk];
                if constexpr (sizeof(T) == 8) acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[q], acc[m], 0, 0, 0);
                else acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[q], acc[m], 0, 0, 0);
            }
        }
        if (live) {
            const T x = X[c];
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = (sizeof(T) == 8) ? 16 * m + lk + 4 * r : 16 * m + 4 * lk + r;
                    W[(i64)p * ldw + c] = (T)0.5 * acc[m][r] + kv[p] * x;
                }
        }
    }
}

// Scatter pass: Y[I][v] = sum over alpha and beta table entries of I.
template <int NV, typename T, bool TILED>
__global__ void __launch_bounds__(256) k_mf_sigma(i64 nd, i64 NB, int nea, int neb, const int2 *__restrict__ ta,
                                                  const int2 *__restrict__ tb, int nvec, const T *__restrict__ W,
                                                  i64 ldw, T *__restrict__ Y) {
    i64 Ia, Ib;
    if (!det_of<TILED>(nd, nd / NB, NB, Ia, Ib)) return;
    const i64 I = Ia * NB + Ib;
    T acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = (T)0;
    const int2 *pa = ta + Ia * nea;
    for (int t = 0; t < nea; ++t) {
        const int2 e = pa[t];
        const T s = (e.y >> 16) ? (T)-1 : (T)1;
        const T *w = W + (i64)(e.y & 0xffff) * ldw + ((i64)e.x * NB + Ib) * nvec;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            if (v < nvec) acc[v] += s * w[v];
    }
    const int2 *pb = tb + Ib * neb;
    for (int t = 0; t < neb; ++t) {
        const int2 e = pb[t];
        const T s = (e.y >> 16) ? (T)-1 : (T)1;
        const T *w = W + (i64)(e.y & 0xffff) * ldw + (Ia * NB + e.x) * nvec;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            if (v < nvec) acc[v] += s * w[v];
    }
    T *y = Y + I * nvec;
#pragma unroll
    for (int v = 0; v < NV; ++v)
        if (v < nvec) y[v] = acc[v];
}

__global__ void k_direct_fill(double *x, i64 n, u64 seed) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        x[i] = (double)(mix64(seed + (u64)i) >> 11) * 0x1p-53 - 0.5;
}

__global__ void k_direct_to_f32(const double *a, float *b, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

#define GIX(p, q, r, s) ((((p) * n + (q)) * n + (r)) * n + (s))

// Diagonal <J|H|J> with the summation order of the CSR build's diagonal.
__global__ void k_mf_diag(i64 nd, i64 NB, const u64 *__restrict__ sa, const u64 *__restrict__ sb, HamView H,
                          double *__restrict__ diag) {
    const i64 J = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= nd) return;
    const int n = H.n;
    const double *one = H.one, *two = H.two;
    const i64 Ja = J / NB;
    const u64 up = sa[Ja], dn = sb[J - Ja * NB];
    double v2 = 0.0;
    for (u64 x = up; x; x &= x - 1) {
        const int ii = __builtin_ctzll(x);
        v2 += one[(n + 1) * ii];
        for (u64 y = x & (x - 1); y; y &= y - 1) {
            const int kk = __builtin_ctzll(y);
            v2 += two[GIX(ii, kk, ii, kk)] - two[GIX(ii, kk, kk, ii)];
        }
        for (u64 y = dn; y; y &= y - 1) {
            const int kk = __builtin_ctzll(y);
            v2 += two[GIX(ii, kk, ii, kk)];
        }
    }
    for (u64 x = dn; x; x &= x - 1) {
        const int ii = __builtin_ctzll(x);
        v2 += one[(n + 1) * ii];
        for (u64 y = x & (x - 1); y; y &= y - 1) {
            const int kk = __builtin_ctzll(y);
            v2 += two[GIX(ii, kk, ii, kk)] - two[GIX(ii, kk, kk, ii)];
        }
    }
    diag[J] = v2;
}

__global__ void k_wallclock(unsigned long long *t) {
    if (threadIdx.x == 0) t[0] = wall_clock64();
}

void mf_diag_launch(i64 nd, i64 NB, const u64 *sa, const u64 *sb, HamView H, double *diag, hipStream_t s) {
    k_mf_diag<<<grid_for(nd, 256), 256, 0, s>>>(nd, NB, sa, sb, H, diag);
    PE_LAUNCH_CHECK();
}

void mf_wallclock_launch(unsigned long long *t, hipStream_t s) {
    k_wallclock<<<1, 64, 0, s>>>(t);
    PE_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
std::vector<u64> colex_strings(int n, int k) {
    std::vector<u64> s;
    if (k == 0) {
        s.push_back(0);
        return s;
    }
    const u64 last = ((1ULL << k) - 1ULL) << (n - k);
    u64 v = (1ULL << k) - 1ULL;
    for (;;) {
        s.push_back(v);
        if (v == last) break;
        const u64 t = v | (v - 1ULL);  // next integer with the same popcount (colex successor)
        v = (t + 1ULL) | (((~t & (t + 1ULL)) - 1ULL) >> (__builtin_ctzll(v) + 1));
    }
    return s;
}

i64 rank_host(u64 det, const std::vector<i64> &binom) {
    i64 r = 0;
    int k = 0;
    while (det) {
        const int p = __builtin_ctzll(det);
        ++k;
        r += binom[(size_t)p * 65 + k];
        det &= det - 1;
    }
    return r;
}

std::vector<int2> string_table(int n, const std::vector<u64> &S, int ne, const std::vector<i64> &binom) {
    std::vector<int2> t(S.size() * (size_t)ne);
    std::vector<std::pair<int, int2>> row;
    for (size_t r = 0; r < S.size(); ++r) {
        const u64 s = S[r];
        row.clear();
        for (int k = 0; k < n; ++k) {
            if (!((s >> k) & 1ULL)) continue;
            for (int l = 0; l < n; ++l) {
                if (l != k && ((s >> l) & 1ULL)) continue;
                u64 K = s;
                int neg = 0;
                if (l != k) {
                    K = s ^ (1ULL << k) ^ (1ULL << l);
                    const int lo = std::min(k, l), hi = std::max(k, l);
                    const u64 between = ((hi >= 64) ? ~0ULL : ((1ULL << hi) - 1ULL)) & ~((1ULL << (lo + 1)) - 1ULL);
                    neg = __builtin_popcountll(s & between) & 1;
                }
                const int i = std::max(k, l), j = std::min(k, l);
                const int p = i * (i + 1) / 2 + j;
                row.push_back({p, int2{(int)rank_host(K, binom), p | (neg << 16)}});
            }
        }
        std::sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (int e = 0; e < ne; ++e) t[r * ne + e] = row[(size_t)e].second;
    }
    return t;
}

// (ij|kl) over packed pairs (zero padded to kp x kp) and k_ij = h_ij - 1/2 sum_r (ir|rj).
void pair_integrals(int n, const double *one, const double *two, int kp, std::vector<double> &V,
                    std::vector<double> &kv) {
    V.assign((size_t)kp * kp, 0.0);
    kv.assign((size_t)kp, 0.0);
    auto g = [&](int p, int q, int r, int t) { return two[(((size_t)p * n + q) * n + r) * n + t]; };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const int p = i * (i + 1) / 2 + j;
            double kk = one[(size_t)i * n + j];
            for (int r = 0; r < n; ++r) kk -= 0.5 * g(i, r, r, j);  // (ir|rj) = <ir|rj>
            kv[(size_t)p] = kk;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l <= k; ++l) V[(size_t)p * kp + k * (k + 1) / 2 + l] = g(i, k, j, l);  // (ij|kl)
        }
}

bool direct_applicable(const pyci_ham *ham, const pyci_wfn *wfn, i64 nrow, i64 ncol, int symmetric) {
    if (wfn->kind != PYCI_WFN_FULLCI || !wfn->colex || !symmetric) return false;
    if (nrow != wfn->ndet || ncol != wfn->ndet || wfn->ndet < 2) return false;
    const i64 n = ham->nbasis;
    if (n > 15) return false;  // pair dimension <= 128: MT <= 8 row tiles
    if (wfn->nup == 0 || wfn->ndn == 0) return false;
    return true;
}

void direct_setup(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn) {
    PE_REQUIRE(direct_applicable(ham, wfn, op->nrow, op->ncol, op->symmetric), PYCI_EUNSUPPORTED,
               "the matrix-free operator needs the full FullCI space in colex order, a square symmetric operator, "
               "nbasis <= 15 and both spins occupied");
    hipStream_t s = op->ctx->stream;
    auto D = std::make_unique<DirectOp>();
    const int n = (int)ham->nbasis;
    D->n = n;
    D->nup = (int)wfn->nup;
    D->ndn = (int)wfn->ndn;
    D->ecore = ham->ecore;
    D->npair = n * (n + 1) / 2;
    D->kpad = (D->npair + 15) / 16 * 16;
    std::vector<i64> binom(65 * 65);
    for (int a = 0; a <= 64; ++a)
        for (int b = 0; b <= 64; ++b) binom[(size_t)a * 65 + b] = binom_host(a, b);
    const std::vector<u64> sa = colex_strings(n, D->nup), sb = colex_strings(n, D->ndn);
    D->na = (i64)sa.size();
    D->nbs = (i64)sb.size();
    PE_REQUIRE(D->na * D->nbs == wfn->ndet, PYCI_EINVAL, "wave function is not the full FullCI space");
    D->nea = D->nup * (n - D->nup + 1);
    D->neb = D->ndn * (n - D->ndn + 1);
    const std::vector<int2> ta = string_table(n, sa, D->nea, binom), tb = string_table(n, sb, D->neb, binom);
    D->tab_a.alloc(ta.size());
    D->tab_b.alloc(tb.size());
    D->str_a.alloc(sa.size());
    D->str_b.alloc(sb.size());
    PE_HIP(hipMemcpyAsync(D->tab_a.p, ta.data(), sizeof(int2) * ta.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->tab_b.p, tb.data(), sizeof(int2) * tb.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->str_a.p, sa.data(), sizeof(u64) * sa.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->str_b.p, sb.data(), sizeof(u64) * sb.size(), hipMemcpyHostToDevice, s));
    // integrals: host copies for the pair matrices, device copies for the diagonal and a later CSR
    const size_t n2 = (size_t)n * n, n4 = n2 * n2;
    std::vector<double> one(n2), two(n4);
    PE_HIP(hipMemcpyAsync(one.data(), ham->one_mo.p, sizeof(double) * n2, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(two.data(), ham->two_mo.p, sizeof(double) * n4, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    D->one_mo.alloc(n2);
    D->two_mo.alloc(n4);
    PE_HIP(hipMemcpyAsync(D->one_mo.p, one.data(), sizeof(double) * n2, hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->two_mo.p, two.data(), sizeof(double) * n4, hipMemcpyHostToDevice, s));
    const int KP = D->kpad;
    std::vector<double> V, kv;
    pair_integrals(n, one.data(), two.data(), KP, V, kv);
    std::vector<float> V32(V.begin(), V.end()), kv32(kv.begin(), kv.end());
    D->V64.alloc(V.size());
    D->k64.alloc(kv.size());
    D->V32.alloc(V.size());
    D->k32.alloc(kv.size());
    PE_HIP(hipMemcpyAsync(D->V64.p, V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->k64.p, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->V32.p, V32.data(), sizeof(float) * V32.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(D->k32.p, kv32.data(), sizeof(float) * kv32.size(), hipMemcpyHostToDevice, s));
    // diagonal (Davidson preconditioner, nrow == 1 solve, diagonal())
    const i64 nd = wfn->ndet;
    op->diag.alloc((size_t)nd);
    HamView H{n, D->one_mo.p, D->two_mo.p};
    k_mf_diag<<<grid_for(nd, 256), 256, 0, s>>>(nd, D->nbs, D->str_a.p, D->str_b.p, H, op->diag.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipStreamSynchronize(s));
    // reference-format size of the full-space CSR: every single and double
    // excitation is present, so each row has the same count
    const i64 na_ = D->nup, va = n - D->nup, nb_ = D->ndn, vb = n - D->ndn;
    const i64 per_row = na_ * va + nb_ * vb + (na_ * (na_ - 1) / 2) * (va * (va - 1) / 2) +
                        (nb_ * (nb_ - 1) / 2) * (vb * (vb - 1) / 2) + na_ * va * nb_ * vb + 1;
    op->size_ref = (per_row - 1) * nd / 2 + nd;
    op->nnz = 0;
    op->direct = std::move(D);
}

template <typename T>
static T *direct_workspace(DirectOp *D, i64 ldw, hipStream_t s);

template <>
double *direct_workspace<double>(DirectOp *D, i64 ldw, hipStream_t s) {
    if (D->ldw64 < ldw) {
        D->work64.alloc((size_t)D->kpad * ldw);
        PE_HIP(hipMemsetAsync(D->work64.p, 0, sizeof(double) * D->kpad * ldw, s));  // pad rows / columns stay zero
        D->ldw64 = ldw;
    }
    return D->work64.p;
}

template <>
float *direct_workspace<float>(DirectOp *D, i64 ldw, hipStream_t s) {
    if (D->ldw32 < ldw) {
        D->work32.alloc((size_t)D->kpad * ldw);
        PE_HIP(hipMemsetAsync(D->work32.p, 0, sizeof(float) * D->kpad * ldw, s));
        D->ldw32 = ldw;
    }
    return D->work32.p;
}

template <int MT, typename T>
static void launch_gemm(i64 ncol, const T *V, const T *kv, const T *X, T *W, i64 ldw, hipStream_t s) {
    // only the column tiles that hold the ncol live columns (ldw >= their end)
    constexpr size_t lds = (size_t)(MT * 16) * (MT * 16 + 1) * sizeof(T);
    const i64 ntiles = (ncol + 63) / 64;
    if (lds <= 65536) {
        // persistent: as many blocks as fit on the 256 CUs (160 KB of LDS each), V loaded once per block
        const i64 per_cu = std::max<i64>(1, std::min<i64>(8, (160 * 1024) / (i64)lds));
        const unsigned grid = (unsigned)std::min<i64>(ntiles, 256 * per_cu);
        k_mf_gemm2<MT, T, true><<<grid, 256, lds, s>>>(ncol, V, kv, X, W, ldw);
    } else {
        k_mf_gemm2<MT, T, false><<<(unsigned)ntiles, 256, 0, s>>>(ncol, V, kv, X, W, ldw);
    }
    PE_LAUNCH_CHECK();
}

// ev (optional): 4 events recorded before, between and after the three passes.
template <int NV, typename T>
static void launch_passes(const DirectOp *D, i64 nd, int nvec, const T *X, T *W, i64 ldw, T *Y, const T *V,
                          const T *kv, hipStream_t s, hipEvent_t *ev) {
    const bool tiled = (D->nbs + 15) / 16 >= 16;  // enough column tiles to spread over the XCDs
    const unsigned g = tiled ? tiled_grid(D->na, D->nbs) : grid_for(nd, 256);
    if (ev) PE_HIP(hipEventRecord(ev[0], s));
    if (tiled)
        k_mf_d<NV, T, true><<<g, 256, 0, s>>>(nd, D->nbs, D->npair, D->nea, D->neb, D->tab_a.p, D->tab_b.p, nvec, X,
                                              W, ldw);
    else
        k_mf_d<NV, T, false><<<g, 256, 0, s>>>(nd, D->nbs, D->npair, D->nea, D->neb, D->tab_a.p, D->tab_b.p, nvec, X,
                                               W, ldw);
    PE_LAUNCH_CHECK();
    if (ev) PE_HIP(hipEventRecord(ev[1], s));
    const i64 ncol = nd * nvec;
    switch (D->kpad / 16) {
        case 1: launch_gemm<1, T>(ncol, V, kv, X, W, ldw, s); break;
        case 2: launch_gemm<2, T>(ncol, V, kv, X, W, ldw, s); break;
        case 3: launch_gemm<3, T>(ncol, V, kv, X, W, ldw, s); break;
        case 4: launch_gemm<4, T>(ncol, V, kv, X, W, ldw, s); break;
        case 5: launch_gemm<5, T>(ncol, V, kv, X, W, ldw, s); break;
        case 6: launch_gemm<6, T>(ncol, V, kv, X, W, ldw, s); break;
        case 7: launch_gemm<7, T>(ncol, V, kv, X, W, ldw, s); break;
        case 8: launch_gemm<8, T>(ncol, V, kv, X, W, ldw, s); break;
        default: throw EngineError(PYCI_EUNSUPPORTED, "pair dimension too large for the matrix-free operator");
    }
    if (ev) PE_HIP(hipEventRecord(ev[2], s));
    if (tiled)
        k_mf_sigma<NV, T, true><<<g, 256, 0, s>>>(nd, D->nbs, D->nea, D->neb, D->tab_a.p, D->tab_b.p, nvec, W, ldw,
                                                  Y);
    else
        k_mf_sigma<NV, T, false><<<g, 256, 0, s>>>(nd, D->nbs, D->nea, D->neb, D->tab_a.p, D->tab_b.p, nvec, W, ldw,
                                                   Y);
    PE_LAUNCH_CHECK();
    if (ev) PE_HIP(hipEventRecord(ev[3], s));
}

template <typename T>
void direct_sigma(const pyci_op *op, int nvec, const T *X, T *Y, hipStream_t s, hipEvent_t *ev) {
    DirectOp *D = op->direct.get();
    PE_REQUIRE(nvec >= 1 && nvec <= 16, PYCI_EUNSUPPORTED, "the matrix-free operator takes 1..16 vectors");
    const i64 nd = op->nrow;
    const i64 ldw = (nd * nvec + 127) / 128 * 128;
    T *W = direct_workspace<T>(D, ldw, s);
    const T *V = std::is_same<T, double>::value ? (const T *)D->V64.p : (const T *)D->V32.p;
    const T *kv = std::is_same<T, double>::value ? (const T *)D->k64.p : (const T *)D->k32.p;
    const i64 ldw_used = std::is_same<T, double>::value ? D->ldw64 : D->ldw32;
    if (nvec == 1) launch_passes<1, T>(D, nd, nvec, X, W, ldw_used, Y, V, kv, s, ev);
    else if (nvec == 2) launch_passes<2, T>(D, nd, nvec, X, W, ldw_used, Y, V, kv, s, ev);
    else if (nvec <= 4) launch_passes<4, T>(D, nd, nvec, X, W, ldw_used, Y, V, kv, s, ev);
    else if (nvec <= 8) launch_passes<8, T>(D, nd, nvec, X, W, ldw_used, Y, V, kv, s, ev);
    else launch_passes<16, T>(D, nd, nvec, X, W, ldw_used, Y, V, kv, s, ev);
}

template void direct_sigma<double>(const pyci_op *, int, const double *, double *, hipStream_t, hipEvent_t *);
template void direct_sigma<float>(const pyci_op *, int, const float *, float *, hipStream_t, hipEvent_t *);

// Per-pass timing of nrep matrix-free sigmas on random X (see pyci_op_bench_sigma_parts).
// Algorithmic bytes and flops per launch, with N = nd * nvec columns and T the value size:
//   gather pass:  reads X (T N), writes D (T npair N)
//   product:      reads D and X, writes E (T (2 npair + 1) N); 2 npair^2 N flops
//   scatter pass: reads E (T npair N), writes Y (T N)
void direct_bench_parts(pyci_op *op, i64 nvec, i64 nrep, int fp32, float *ms, double *bytes, double *flops) {
    DirectOp *D = op->direct.get();
    hipStream_t s = op->ctx->stream;
    const i64 nd = op->nrow, N = nd * nvec;
    DevBuf<double> x((size_t)N), y((size_t)N);
    DevBuf<float> x32, y32;
    k_direct_fill<<<1024, 256, 0, s>>>(x.p, N, 12345);
    PE_LAUNCH_CHECK();
    if (fp32) {
        x32.alloc((size_t)N);
        y32.alloc((size_t)N);
        k_direct_to_f32<<<1024, 256, 0, s>>>(x.p, x32.p, N);
        PE_LAUNCH_CHECK();
    }
    std::vector<hipEvent_t> ev((size_t)(4 * nrep));
    for (auto &e : ev) PE_HIP(hipEventCreate(&e));
    auto run = [&](hipEvent_t *e) {
        if (fp32) direct_sigma<float>(op, (int)nvec, x32.p, y32.p, s, e);
        else direct_sigma<double>(op, (int)nvec, x.p, y.p, s, e);
    };
    try {
        run(nullptr);  // warm-up (workspace allocation)
        for (i64 r = 0; r < nrep; ++r) run(&ev[(size_t)(4 * r)]);
        PE_HIP(hipEventSynchronize(ev.back()));
        double acc[3] = {0, 0, 0};
        for (i64 r = 0; r < nrep; ++r)
            for (int p = 0; p < 3; ++p) {
                float t = 0.f;
                PE_HIP(hipEventElapsedTime(&t, ev[(size_t)(4 * r + p)], ev[(size_t)(4 * r + p + 1)]));
                acc[p] += t;
            }
        const double T = fp32 ? 4.0 : 8.0, P = (double)D->npair, NN = (double)N;
        for (int p = 0; p < 3; ++p) ms[p] = (float)(acc[p] / (double)nrep);
        bytes[0] = T * NN * (1.0 + P);
        bytes[1] = T * NN * (2.0 * P + 1.0);
        bytes[2] = T * NN * (P + 1.0);
        flops[0] = 2.0 * (double)(D->nea + D->neb) * NN;
        flops[1] = 2.0 * P * P * NN;
        flops[2] = 2.0 * (double)(D->nea + D->neb) * NN;
    } catch (...) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
}

// Materialises the CSR of a matrix-free operator (for exports and element
// queries), from a regenerated full-space wave function and the stored integrals.
void direct_ensure_csr(pyci_op *op) {
    DirectOp *D = op->direct.get();
    if (!D || D->csr_ready) return;
    pyci_ctx *ctx = op->ctx;
    pyci_ham ham;
    ham.ctx = ctx;
    ham.nbasis = D->n;
    ham.ecore = D->ecore;
    ham.one_mo.alloc((size_t)D->n * D->n);
    ham.two_mo.alloc((size_t)D->n * D->n * D->n * D->n);
    hipStream_t s = ctx->stream;
    PE_HIP(hipMemcpyAsync(ham.one_mo.p, D->one_mo.p, sizeof(double) * D->n * D->n, hipMemcpyDeviceToDevice, s));
    PE_HIP(hipMemcpyAsync(ham.two_mo.p, D->two_mo.p, sizeof(double) * D->n * D->n * D->n * D->n,
                          hipMemcpyDeviceToDevice, s));
    pyci_wfn wfn;
    wfn.ctx = ctx;
    wfn.kind = PYCI_WFN_FULLCI;
    wfn.nbasis = D->n;
    wfn.nup = D->nup;
    wfn.ndn = D->ndn;
    wfn.nkey = 2;
    wfn.nsd = D->nbs;
    wfn.ndet = op->nrow;
    wfn.dets.alloc((size_t)(2 * op->nrow));
    fill_all_dets_device(ctx, PYCI_WFN_FULLCI, D->n, D->nup, D->ndn, wfn.dets.p, 1);
    wfn_build_index(&wfn);
    DevBuf<i64> rp, lc;
    DevBuf<int32_t> cols;
    DevBuf<double> vals, dg;
    i64 nnz = 0, low = 0;
    build_rows(op, &ham, &wfn, 0, op->nrow, op->ncol, rp, cols, vals, lc, dg, nnz, low);
    PE_HIP(hipStreamSynchronize(s));
    op->rowptr = std::move(rp);
    op->lowcnt = std::move(lc);
    op->cols = std::move(cols);
    op->vals = std::move(vals);
    op->vals32.release();
    op->nnz = nnz;
    op->size_ref = low;
    op->segments.assign(1, {op->nrow, op->ncol});
    D->csr_ready = true;
}

}  // namespace pe
