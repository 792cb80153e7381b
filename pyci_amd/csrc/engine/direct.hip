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

#include <mutex>

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

__device__ inline f64x4 mfma16(double a, double b, f64x4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f32x4 mfma16(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// An empty asm statement that takes v as an operand: the value must exist here,
// so the load that produces it is issued unconditionally where the source has it
// (instead of being sunk into a branch of its own that waits for it alone).
__device__ inline void pin(int &v) { asm volatile("" : "+v"(v)); }
__device__ inline void pin(float &v) { asm volatile("" : "+v"(v)); }
__device__ inline void pin(double &v) { asm volatile("" : "+v"(v)); }

// Workspace layout: column c = J * nvec + v (determinant J, vector v) of pair row
// p sits at W[(c / 64) * KP * 64 + p * 64 + c % 64]: the rows of one 64-column tile
// form one contiguous block of KP * 64 values, which the product pass writes and
// the scatter pass reads in 512-byte (FP64) runs.
__host__ __device__ inline i64 wt_at(i64 c, int p, int KP) { return (c >> 6) * (i64)KP * 64 + p * 64 + (c & 63); }

// Gather + product pass, fused. Per 64-column tile:
//   D[q][c] = sum over the alpha and beta table entries of J with pair q of s X[K][v]
//   E[p][c] = 1/2 sum_q V[p][q] D[q][c] + k[p] X[c]        -> W (pair rows; pad rows are
//                                                             never read)
// 768 threads: waves 4..11 gather, waves 0..3 multiply; the grid is persistent
// (tile = blockIdx.x + step * gridDim.x) and D never reaches HBM. With two LDS
// buffers (SB false) the gather of tile t+1 runs beside the product of tile t, one
// barrier per tile. With one buffer (SB true, where two do not fit beside V) the
// gather waves add tile t into the buffer between two barriers and load the X
// values of tile t+1 into registers while the product of tile t runs.
// Gather: lane = column; wave 4 + g takes table entries g, g + 8, ... (PE slots,
// 8 PE >= nea + neb) of the merged table (alpha rows, then the beta entries slot by
// slot, then one sentinel entry that every empty slot reads: it adds into the last
// buffer row, which nothing reads), loading the entries of its next tile while X values are in
// flight. Indices are 32-bit (nd nvec < 2^31). A (pair, column) cell receives at
// most one alpha and one beta term (a string's entries have distinct pairs), added
// onto zero, so the LDS atomic adds give the same bits in any order. D is held in
// FP64 for both value types (LDS FP64 adds are fast where FP32 ones are not; an
// FP32 D value is the sum of two FP32 terms, rounded once).
// Product: wave w owns columns 16 w .. 16 w + 15 of the tile and the row tiles; the
// k loop is unrolled, with the A and B fragments of step q + 1 loaded before the
// MFMAs of step q; each lane clears the D values it has read, so the buffer is zero
// again when it is next filled. MFMA 16x16x4: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator r of lane l is D[4 (l / 16) + r][l % 16] for the
// FP32 instruction and D[l / 16 + 4 r][l % 16] for the FP64 one (lane layouts checked
// on the device). V (kpad x kpad, zero padded) is read from LDS (VM_FULL, row
// stride VC + 1) or through the caches (VM_GLOBAL). For kpad >= 96 the pair count
// is fixed by kpad (91, 105, 120 for nbasis 13, 14, 15), so the k extent VC, the
// buffer rows and the row tiles stop at the last pair. Buffer row stride ld = 80
// puts the four 16-lane groups of a B read on distinct LDS banks (64 where 80 does
// not fit).
constexpr int kDprodThreads = 768;
enum { VM_GLOBAL = 0, VM_FULL = 1 };

// pair count implied by a pair dimension of kpad = 16 MT >= 96 (else 0: not fixed)
__host__ __device__ constexpr int fixed_npair(int MT) { return MT == 6 ? 91 : MT == 7 ? 105 : MT == 8 ? 120 : 0; }
// row tiles of the fused pass that hold a pair, for a pair dimension of 16 MT
__host__ __device__ constexpr int pair_tiles(int MT) { return fixed_npair(MT) ? (fixed_npair(MT) + 15) / 16 : MT; }

// k extent of the product (a multiple of 4); the sink row of the gather buffers is row VC
__host__ __device__ constexpr int dprod_vc(int MT) {
    return fixed_npair(MT) ? 4 * ((fixed_npair(MT) + 3) / 4) : 16 * MT;
}

int mf_sink_row(int kpad) { return dprod_vc(kpad / 16); }

struct DprodArgs {
    int ncol, NB, nvec, npair, nea, ntot;
    int vstride, v0;  // X[J][v] of vector v0 + v (v < nvec) at X[J * vstride + v0 + v]
    int colmajor;     // or, with colmajor, at X[v * nd + J] (vectors contiguous)
    int nd;
    const long long *dyn;  // batched solver: device {vectors (0: nothing), X offset, Y offset}, or null
    int offb;  // start of the beta entries in the merged table, NA * nea (entry t of beta string r at offb + t NB + r)
    int sent;  // index of the sentinel entry
    int ld;
    // live-row layout (ML > 0): column c = Ja * nbp + Jb * nvec + v (nbp: NB nvec rounded
    // up to 64, so a tile holds one alpha string; columns with Jb >= NB are empty), output
    // row r of a tile of alpha string Ka is pair live[Ka * 16 ML + r] (r < nl; pad rows
    // name a zero row of V)
    int nbp, nl;
    const int *live;
};

template <int MT, typename T, bool SB, int VM, int PE, int ML>
__global__ void __launch_bounds__(768) k_mf_dprod(DprodArgs A, const int2 *__restrict__ tab, const T *__restrict__ X0,
                                                  const T *__restrict__ V, const T *__restrict__ kv,
                                                  T *__restrict__ W) {
    const T *__restrict__ X = X0;
    if (A.dyn) {
        A.nvec = (int)A.dyn[0];
        A.ncol = A.nd * A.nvec;
        X += A.dyn[1];
        if (A.ncol == 0) return;  // every block: nothing to multiply this step
    }
    constexpr int KP = MT * 16, NT = kDprodThreads;
    constexpr int NP = fixed_npair(MT);             // 0: pair count known only at run time
    constexpr int VC = dprod_vc(MT), NQ = VC / 4;   // k extent and k steps
    constexpr int MTL = NP ? (NP + 15) / 16 : MT;   // row tiles that hold a pair
    constexpr int MR = ML ? ML : MTL;               // output row tiles
    constexpr int RS = ML ? 16 * ML : KP;           // rows of a tile's block of W
    static_assert(ML == 0 || ML < MTL, "the live rows must be fewer than the pair rows");
    constexpr int BR = VC + 1, LS = VC + 1, NBUF = SB ? 1 : 2;
    static_assert(!(SB && VM == VM_GLOBAL), "one buffer only with V in LDS");
    // With V in LDS and a free padding row VC - 1 in the k extent (every pair
    // dimension but 120), the whole update E = V D / 2 + k x is the product: the LDS
    // copy of V is scaled by 1/2 (exact) and holds k in column VC - 1, and row VC - 1
    // of the B operand is each column's own X value, so the epilogue only stores.
    constexpr bool FOLD = VM == VM_FULL && (NP == 0 || NP < VC);
    extern __shared__ unsigned char smem_raw[];
    double *buf0 = reinterpret_cast<double *>(smem_raw);     // NBUF buffers of BR x ld
    T *ks = reinterpret_cast<T *>(buf0 + NBUF * BR * A.ld);  // k[p], kpad values
    T *vs = ks + KP;                                         // V rows 0..KP-1, columns 0..VC-1
    const int ld = A.ld;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ntiles = (A.ncol + 63) / 64;
    const int nmine = (ntiles > (int)blockIdx.x) ? (ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
    for (int i = threadIdx.x; i < NBUF * BR * ld; i += NT) buf0[i] = 0.0;
    for (int i = threadIdx.x; i < KP; i += NT) ks[i] = kv[i];
    if constexpr (VM == VM_FULL)
        for (int i = threadIdx.x; i < KP * VC; i += NT) {
            const int r = i / VC, q = i % VC;
            vs[r * LS + q] = !FOLD ? V[r * KP + q] : (q == VC - 1) ? kv[r] : (T)0.5 * V[r * KP + q];
        }
    // The two roles run separate loops with the same barrier count, so that neither
    // role's registers are live in the other.
    if (wave >= 4) {
        const int g = wave - 4, sa = A.colmajor ? A.NB : A.NB * A.vstride, sb = A.colmajor ? 1 : A.vstride;
        // loads the table entries of the tile of `step` into nx and its X bases: the X
        // index of an alpha entry is en.x * sa + na, of a beta entry en.x * sb + nb
        auto fetch = [&](int step, int2 *nx, int &na, int &nb) {
            const int c0 = ((int)blockIdx.x + step * (int)gridDim.x) * 64 + lane;
            bool live = step < nmine && c0 < A.ncol;
            const unsigned c = live ? (unsigned)c0 : 0u;
            unsigned v, Ja, Jb;
            if constexpr (ML > 0) {
                Ja = c / (unsigned)A.nbp;
                const unsigned r = c - Ja * (unsigned)A.nbp;
                Jb = r / (unsigned)A.nvec;
                v = r - Jb * (unsigned)A.nvec;
                if (Jb >= (unsigned)A.NB) {  // empty column of the padded layout
                    live = false;
                    Ja = Jb = v = 0;
                }
            } else {
                const unsigned J = c / (unsigned)A.nvec;
                v = c - J * (unsigned)A.nvec;
                Ja = J / (unsigned)A.NB;
                Jb = J - Ja * (unsigned)A.NB;
            }
            // alpha rows string-major; beta entries slot-major (entry t of every beta
            // string contiguous, so the lanes' loads of one slot are one run)
            const int ra = (int)(Ja * (unsigned)A.nea), rb = A.offb + (int)Jb - A.nea * A.NB;
            if (A.colmajor) {
                na = (int)(v * (unsigned)A.nd + Jb);
                nb = (int)(v * (unsigned)A.nd + Ja * (unsigned)A.NB);
            } else {
                na = (int)(Jb * (unsigned)A.vstride + v) + A.v0;
                nb = (int)(Ja * (unsigned)sa + v) + A.v0;
            }
#pragma unroll
            for (int u = 0; u < PE; ++u) {
                const int t = g + 8 * u;
                const int at = (t < A.nea) ? ra + t : rb + t * A.NB;
                nx[u] = tab[(live && t < A.ntot) ? at : A.sent];
            }
        };
        auto xload = [&](const int2 *en, int xa, int xb, T *val) {
#pragma unroll
            for (int u = 0; u < PE; ++u) {
                const int t = g + 8 * u;
                val[u] = X[(t < A.nea) ? en[u].x * sa + xa : en[u].x * sb + xb];
            }
        };
        auto add = [&](const int2 *en, T *val, double *buf) {
#pragma unroll
            for (int u = 0; u < PE; ++u) pin(val[u]);
#pragma unroll
            for (int u = 0; u < PE; ++u) {
                const double x = (double)val[u];
                atomicAdd(buf + lane + (en[u].y & 0xffff) * ld, (en[u].y >> 16) ? -x : x);
            }
        };
        int2 en[PE], nx[PE];
        T val[PE];
        int xa, xb, na, nb;
        fetch(0, en, xa, xb);
        if constexpr (!SB) {
            __syncthreads();
            for (int step = 0; step <= nmine; ++step) {
                if (step < nmine) {
                    xload(en, xa, xb, val);
                    fetch(step + 1, nx, na, nb);  // in flight beside the X loads
                    add(en, val, buf0 + (step & 1) * BR * ld);
#pragma unroll
                    for (int u = 0; u < PE; ++u) en[u] = nx[u];
                    xa = na;
                    xb = nb;
                }
                __syncthreads();
            }
        } else {
            xload(en, xa, xb, val);
            fetch(1, nx, na, nb);
            __syncthreads();
            for (int step = 0; step < nmine; ++step) {
                __syncthreads();  // the product waves are done with the buffer
                add(en, val, buf0);
                __syncthreads();  // the buffer holds D of this tile
#pragma unroll
                for (int u = 0; u < PE; ++u) en[u] = nx[u];
                xload(en, na, nb, val);  // the next tile's X values, in flight during the product
                fetch(step + 2, nx, na, nb);
            }
        }
    } else {
        const int lr = lane & 15, lk = lane >> 4;
        typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type acc_t;
        // pair of A row 16 m + lr: the row itself, or (live rows) the pair of live row
        // 16 m + lr of the tile's alpha string, loaded one tile ahead with x
        int prow[ML ? ML : 1], pnext[ML ? ML : 1];
        auto rows_of = [&](int tile, int *r) {
            if constexpr (ML > 0) {
                const int Ka = min(tile, ntiles - 1) * 64 / A.nbp;
#pragma unroll
                for (int m = 0; m < ML; ++m) r[m] = A.live[Ka * RS + 16 * m + lr];
            }
        };
        auto arow = [&](int m) -> int {
            if constexpr (ML > 0) return prow[m];
            else return 16 * m + lr;
        };
        auto aload = [&](int m, int q) -> T {
            if constexpr (VM == VM_FULL) return vs[arow(m) * LS + 4 * q + lk];
            else return V[arow(m) * KP + 4 * q + lk];
        };
        // X of column c (determinant c / nvec, vector c % nvec; live rows: determinant
        // (c / nbp, (c % nbp) / nvec), empty columns clamped), loaded one tile ahead
        auto xat = [&](int c) -> T {
            const int cc = min(c, A.ncol - 1);
            if constexpr (ML > 0) {
                const int Ja = cc / A.nbp, r = cc - Ja * A.nbp, q = r / A.nvec, v = r - q * A.nvec;
                return X[(Ja * A.NB + min(q, A.NB - 1)) * A.vstride + A.v0 + v];
            }
            const int J = cc / A.nvec, v = cc - J * A.nvec;
            return A.colmajor ? X[v * A.nd + J] : X[J * A.vstride + A.v0 + v];
        };
        T x = xat((int)blockIdx.x * 64 + 16 * wave + lr);
        rows_of((int)blockIdx.x, pnext);
        const int nlim = ML ? A.nl : A.npair;  // rows that are stored
        // Two-buffer plans with the folded epilogue store a tile's E values during the
        // next tile's k loop, one per k step after its MFMAs (pv, pw, pok: the values,
        // their column and whether it exists), so no run of stores stands between the
        // MFMAs of consecutive tiles; the last tile's values are stored after the loop.
        // Rows past the last pair are not stored (no pass reads them).
        constexpr bool PIPE = FOLD && !SB;
        T pv[MR][4];
        T *pw = W;
        bool pok = false;
        auto row_of = [&](int m, int r) { return (sizeof(T) == 8) ? 16 * m + lk + 4 * r : 16 * m + 4 * lk + r; };
        auto product = [&](int tile, double *buf) {
            const int c = tile * 64 + 16 * wave + lr;
            double *bb = buf + 16 * wave + lr;
            if constexpr (ML > 0)
#pragma unroll
                for (int m = 0; m < ML; ++m) prow[m] = pnext[m];
            acc_t acc[MR];
#pragma unroll
            for (int m = 0; m < MR; ++m) acc[m] = acc_t{};
            if constexpr (VM == VM_GLOBAL) {
                // V through the caches (nbasis 15 in FP64 only): one k step at a time
#pragma unroll 1
                for (int q = 0; q < NQ; ++q) {
                    const T bq = (T)bb[(4 * q + lk) * ld];
                    bb[(4 * q + lk) * ld] = 0.0;
#pragma unroll
                    for (int m = 0; m < MR; ++m) acc[m] = mfma16(aload(m, q), bq, acc[m]);
                }
            } else {
                T ac[MR], an[MR], bc, bn;
                bc = (T)bb[lk * ld];
                bb[lk * ld] = 0.0;
                if (FOLD && NQ == 1 && lk == 3) bc = x;
#pragma unroll
                for (int m = 0; m < MR; ++m) ac[m] = aload(m, 0);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    if (q + 1 < NQ) {
                        bn = (T)bb[(4 * q + 4 + lk) * ld];
                        bb[(4 * q + 4 + lk) * ld] = 0.0;
                        if (FOLD && q + 2 == NQ && lk == 3) bn = x;  // row VC - 1
#pragma unroll
                        for (int m = 0; m < MR; ++m) an[m] = aload(m, q + 1);
                    }
                    // the loads of step q + 1 go out before the MFMAs of step q, and the
                    // loads of later steps after them (no instruction crosses either point;
                    // not in the one-buffer plans, whose register budget it would exceed)
                    if constexpr (!SB) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int m = 0; m < MR; ++m) acc[m] = mfma16(ac[m], bc, acc[m]);
                    if constexpr (!SB) __builtin_amdgcn_sched_barrier(0);
                    if constexpr (PIPE)
                        if (q < MR * 4 && pok && row_of(q / 4, q % 4) < nlim)
                            pw[row_of(q / 4, q % 4) * 64] = pv[q / 4][q % 4];
                    asm volatile("" ::: "memory");
#pragma unroll
                    for (int m = 0; m < MR; ++m) ac[m] = an[m];
                    bc = bn;
                }
            }
            if constexpr (PIPE) {
                // the previous tile's values beyond the k steps, then this tile's in their place
#pragma unroll
                for (int j = NQ; j < MR * 4; ++j)
                    if (pok && row_of(j / 4, j % 4) < nlim) pw[row_of(j / 4, j % 4) * 64] = pv[j / 4][j % 4];
#pragma unroll
                for (int m = 0; m < MR; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) pv[m][r] = acc[m][r];
                pw = W + (i64)tile * RS * 64 + 16 * wave + lr;
                pok = c < A.ncol;
            } else if (c < A.ncol) {
                T *w = W + (i64)tile * RS * 64 + 16 * wave + lr;
                const int *lrow = ML ? A.live + (tile * 64 / A.nbp) * RS : nullptr;
#pragma unroll
                for (int m = 0; m < MR; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int p = row_of(m, r);
                        if (ML && p >= nlim) continue;
                        if constexpr (FOLD)
                            w[p * 64] = acc[m][r];
                        else
                            w[p * 64] = (T)0.5 * acc[m][r] + ks[ML ? lrow[p] : p] * x;
                    }
            }
            x = xat(c + 64 * (int)gridDim.x);
            rows_of(tile + (int)gridDim.x, pnext);
        };
        if constexpr (!SB) {
            __syncthreads();
            for (int step = 0; step <= nmine; ++step) {
                if (step >= 1)
                    product((int)blockIdx.x + (step - 1) * (int)gridDim.x, buf0 + ((step - 1) & 1) * BR * ld);
                __syncthreads();
            }
            if constexpr (PIPE)
                if (pok)
#pragma unroll
                    for (int m = 0; m < MR; ++m)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (row_of(m, r) < nlim) pw[row_of(m, r) * 64] = pv[m][r];
        } else {
            __syncthreads();
            for (int step = 0; step < nmine; ++step) {
                __syncthreads();
                __syncthreads();
                product((int)blockIdx.x + step * (int)gridDim.x, buf0);
            }
        }
    }
}

// Scatter pass: Y[I][v] = sum over the alpha and beta table entries of I of s E[pair][K v].
// With alpha_only the beta terms are left out (see k_mf_symm). The E value of alpha entry
// (Ka, r) for vector v sits at column Ka * cs + Ib * nvec + v, row r of tiles of KP rows:
// cs = NB nvec and r = the pair (pass cs = 0), or the live-row layout (alpha_only; cs =
// NB nvec rounded up to 64, r = the live row, KP = the live tile rows).
template <int NV, typename T, bool TILED>
__global__ void __launch_bounds__(256) k_mf_sigma(i64 nd, i64 NB, int nea, int neb, int KP,
                                                  const int2 *__restrict__ ta, const int2 *__restrict__ tb, int nvec,
                                                  int vstride, int v0, int alpha_only, int colmajor,
                                                  const long long *dyn, const T *__restrict__ W, T *__restrict__ Y,
                                                  i64 cs) {
    if (dyn) {
        nvec = (int)dyn[0];
        Y += dyn[2];
        if (nvec == 0) return;
    }
    if (cs == 0) cs = NB * nvec;
    i64 Ia, Ib;
    if (!det_of<TILED>(nd, nd / NB, NB, Ia, Ib)) return;
    const i64 I = Ia * NB + Ib;
    T acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = (T)0;
    // CH table entries, then their CH x NV E values, are loaded before any is used:
    // the loads of a chunk are independent, so a wave waits twice per chunk instead
    // of twice per entry. Indices past the row and vectors past nvec are clamped to
    // valid ones and their values dropped.
    constexpr int CH = NV >= 8 ? 2 : 16 / NV;
    const auto terms = [&](const int2 *row, int n, bool alpha) {
        for (int t0 = 0; t0 < n; t0 += CH) {
            int2 e[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) e[u] = row[min(t0 + u, n - 1)];
            T w[CH][NV];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const i64 c = alpha ? (i64)e[u].x * cs + Ib * nvec : (Ia * NB + e[u].x) * nvec;
#pragma unroll
                for (int v = 0; v < NV; ++v) w[u][v] = W[wt_at(c + min(v, nvec - 1), e[u].y & 0xffff, KP)];
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const T sg = (t0 + u >= n) ? (T)0 : (e[u].y >> 16) ? (T)-1 : (T)1;
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] += sg * w[u][v];
            }
        }
    };
    terms(ta + Ia * nea, nea, true);
    if (!alpha_only) terms(tb + Ib * neb, neb, false);
#pragma unroll
    for (int v = 0; v < NV; ++v)
        if (v < nvec) Y[colmajor ? v * nd + I : I * vstride + v0 + v] = acc[v];
}

// Alpha terms of one row-major vector (the first pass of a sector sigma): one wave per
// alpha string Ia and 64 consecutive beta strings, so the table row of Ia is the same
// for the whole wave and comes through scalar loads; each chunk of CH entries issues
// its CH E loads before adding any. Sums in the order of k_mf_sigma (same bits).
// E value of entry (Ka, r) at column Ka * nbp + Ib, row r of tiles of KP rows: the
// pair layout (nbp = NB, r = the pair) or the live-row one (nbp = NB rounded up to 64,
// r = the rank of the pair among the live rows of Ka, KP = the live tile rows).
template <typename T>
__global__ void __launch_bounds__(256) k_mf_alpha(int na, int NB, int nbp, int nchunk, int nea, int KP,
                                                  const int2 *__restrict__ ta, const T *__restrict__ W,
                                                  T *__restrict__ Y) {
    const int wv = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int Ia = __builtin_amdgcn_readfirstlane(wv / nchunk);
    const int q = __builtin_amdgcn_readfirstlane(wv - (wv / nchunk) * nchunk);
    if (Ia >= na) return;
    const int Ib = q * 64 + lane;
    const bool live = Ib < NB;
    const i64 ib = live ? Ib : NB - 1;
    const int2 *row = ta + (i64)Ia * nea;
    constexpr int CH = 24;
    T acc = (T)0;
    for (int t0 = 0; t0 < nea; t0 += CH) {
        int2 e[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) e[u] = row[min(t0 + u, nea - 1)];
        T w[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) w[u] = W[wt_at((i64)e[u].x * nbp + ib, e[u].y & 0xffff, KP)];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const T sg = (t0 + u >= nea) ? (T)0 : (e[u].y >> 16) ? (T)-1 : (T)1;
            acc += sg * w[u];
        }
    }
    if (live) Y[(i64)Ia * NB + Ib] = acc;
}

// Sigma of a vector in a parity sector of the alpha <-> beta exchange P (nup == ndn,
// J = (Ja, Jb) -> (Jb, Ja); x(P J) = s x(J)): D and E then have the same parity,
// E[p][(Ia, K)] = s E[p][(K, Ia)], so the beta terms of determinant I equal s times
// the alpha terms of P(I) (the alpha and beta string tables are the same) and
//   y(I) = A(I) + s A(P I),   A = the alpha terms (k_mf_sigma with alpha_only).
// In place on Y: the thread of I with Ia <= Ib owns the pair (I, P I).
template <typename T>
__global__ void k_mf_symm(T *Y, i64 NB, int nvec, int vstride, int v0, int colmajor, const long long *dyn, T s) {
    if (dyn) {
        nvec = (int)dyn[0];
        Y += dyn[2];
        if (nvec == 0) return;
    }
    const i64 nd = NB * NB;
    const i64 sj = colmajor ? 1 : vstride, sv = colmajor ? nd : 1, o = colmajor ? 0 : v0;
    for (i64 I = (i64)blockIdx.x * blockDim.x + threadIdx.x; I < nd; I += (i64)gridDim.x * blockDim.x) {
        const i64 Ia = I / NB, Ib = I - Ia * NB;
        if (Ia > Ib) continue;
        T *y0 = Y + I * sj + o, *y1 = Y + (Ib * NB + Ia) * sj + o;
        for (int v = 0; v < nvec; ++v) {
            const T a = y0[v * sv], b = y1[v * sv];
            if (Ia == Ib) {
                y0[v * sv] = a + s * a;
            } else {
                y0[v * sv] = a + s * b;
                y1[v * sv] = b + s * a;
            }
        }
    }
}

// k_mf_symm for one row-major vector, by tiles: block b owns the 32 x 32 tile (ti, tj)
// of (alpha, beta) strings with ti <= tj and its mirror (tj, ti), read into LDS with
// coalesced rows and written back the same way (the same sums as k_mf_symm).
template <typename T>
__global__ void __launch_bounds__(256) k_mf_symm_tiles(T *Y, int NB, T s) {
    __shared__ T A[32][33], B[32][33];
    const int b = (int)blockIdx.x;
    int tj = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while (tj > 0 && tj * (tj + 1) / 2 > b) --tj;
    while ((tj + 1) * (tj + 2) / 2 <= b) ++tj;
    const int ti = b - tj * (tj + 1) / 2;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int ia = ti * 32 + r, ib = tj * 32 + tx, ja = tj * 32 + r, jb = ti * 32 + tx;
        A[r][tx] = (ia < NB && ib < NB) ? Y[(i64)ia * NB + ib] : (T)0;
        B[r][tx] = (ja < NB && jb < NB) ? Y[(i64)ja * NB + jb] : (T)0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int ia = ti * 32 + r, ib = tj * 32 + tx, ja = tj * 32 + r, jb = ti * 32 + tx;
        // Y(ia, ib) pairs with Y(ib, ia) = B[tx][r]; Y(ja, jb) with Y(jb, ja) = A[tx][r]
        if (ia < NB && ib < NB) Y[(i64)ia * NB + ib] = A[r][tx] + s * B[tx][r];
        if (ti != tj && ja < NB && jb < NB) Y[(i64)ja * NB + jb] = B[r][tx] + s * A[tx][r];
    }
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
    PE_REQUIRE(n >= 1 && n <= 20, PYCI_EUNSUPPORTED, "string tables are built for nbasis <= 20");
    (void)binom;
    // S holds every string of its popcount in colex (= increasing) order, so the rank
    // of a string is its position in S: a direct table over all 2^n masks
    std::vector<int> rk((size_t)1 << n, -1);
    for (size_t r = 0; r < S.size(); ++r) rk[(size_t)S[r]] = (int)r;
    std::vector<int2> t(S.size() * (size_t)ne);
    // Pairs p = i (i + 1) / 2 + j (i >= j) are visited in increasing order, so each row
    // comes out sorted by pair: (i, i) for i occupied, (i, j) for exactly one of i, j
    // occupied (k the occupied one, l the other).
    for (size_t r = 0; r < S.size(); ++r) {
        const u64 s = S[r];
        int e = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                const int oi = (int)((s >> i) & 1ULL), oj = (int)((s >> j) & 1ULL);
                if (i == j ? !oi : oi == oj) continue;
                u64 K = s;
                int neg = 0;
                if (i != j) {
                    K = s ^ (1ULL << i) ^ (1ULL << j);
                    const u64 between = ((1ULL << i) - 1ULL) & ~((1ULL << (j + 1)) - 1ULL);
                    neg = __builtin_popcountll(s & between) & 1;
                }
                const int p = i * (i + 1) / 2 + j, rank = rk[(size_t)K];
                PE_REQUIRE(e < ne && rank >= 0, PYCI_EINVAL, "string table: the strings are not a full colex list");
                t[r * ne + e++] = int2{rank, p | (neg << 16)};
            }
        PE_REQUIRE(e == ne, PYCI_EINVAL, "string table: unexpected number of entries");
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
    const std::vector<int2> ta = string_table(n, sa, D->nea, binom);
    const std::vector<int2> tb = (D->nup == D->ndn) ? ta : string_table(n, sb, D->neb, binom);  // same strings
    D->tab_a.alloc(ta.size());
    D->tab_b.alloc(tb.size());
    std::vector<int2> merged(ta);
    for (int t = 0; t < D->neb; ++t)
        for (i64 r = 0; r < (i64)sb.size(); ++r) merged.push_back(tb[(size_t)r * D->neb + t]);
    merged.push_back(int2{0, dprod_vc(D->kpad / 16)});  // sentinel: empty slots of the fused pass add into the sink row
    D->tab.alloc(merged.size());
    PE_HIP(hipMemcpyAsync(D->tab.p, merged.data(), sizeof(int2) * merged.size(), hipMemcpyHostToDevice, s));
    // Parity-sector layout (nup == ndn): the alpha scatter reads, at column (Ka, Ib), only
    // pairs {k, l} with one orbital occupied in Ka and the other empty, or k = l occupied:
    // exactly the pairs of Ka's own table row. Their E rows alone are formed and stored,
    // in that row's (pair) order; used where the fewer row tiles outweigh padding the
    // beta strings to 64 (so that a 64-column tile holds one alpha string).
    std::vector<int> live;
    std::vector<int2> tal;
    if (D->nup == D->ndn) {
        const int ntot = 2 * D->nea, ml = ntot <= 64 ? 2 : ntot <= 96 ? 3 : 4;
        const i64 nbp = (D->nbs + 63) / 64 * 64;
        PE_REQUIRE(D->npair < D->kpad, PYCI_EUNSUPPORTED, "no zero row in the pair dimension");
        if (ntot <= 128 && ml < pair_tiles(D->kpad / 16) && nbp * ml < D->nbs * pair_tiles(D->kpad / 16)) {
            D->ml = ml;
            D->nbp = nbp;
            const int rs = 16 * ml;
            live.assign((size_t)D->na * rs, D->kpad - 1);  // pad rows: row kpad - 1 of V and k is zero
            std::vector<int> rank((size_t)D->na * D->kpad, -1);
            for (i64 a = 0; a < D->na; ++a)
                for (int t = 0; t < D->nea; ++t) {
                    const int p = ta[(size_t)(a * D->nea + t)].y & 0xffff;
                    live[(size_t)(a * rs + t)] = p;
                    rank[(size_t)(a * D->kpad + p)] = t;
                }
            tal.resize(ta.size());
            for (i64 a = 0; a < D->na; ++a)
                for (int t = 0; t < D->nea; ++t) {
                    const int2 e = ta[(size_t)(a * D->nea + t)];
                    const int r = rank[(size_t)e.x * D->kpad + (e.y & 0xffff)];
                    PE_REQUIRE(r >= 0, PYCI_EUNSUPPORTED, "string table: pair not among the live rows");
                    tal[(size_t)(a * D->nea + t)] = int2{e.x, r | (e.y & ~0xffff)};
                }
            D->live.alloc(live.size());
            D->tab_al.alloc(tal.size());
            PE_HIP(hipMemcpyAsync(D->live.p, live.data(), sizeof(int) * live.size(), hipMemcpyHostToDevice, s));
            PE_HIP(hipMemcpyAsync(D->tab_al.p, tal.data(), sizeof(int2) * tal.size(), hipMemcpyHostToDevice, s));
        }
    }
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

// Workspace of the E values: ceil(columns / 64) tiles of kpad x 64 (tile-major),
// held by the context and shared by its operators (their sigmas run in order on
// the context's stream); grown on demand, never cleared (the scatter pass reads
// only what the product pass of the same sigma wrote).
template <typename T>
static T *direct_workspace(const pyci_op *op, i64 elems, DevBuf<unsigned char> *ws) {
    DevBuf<unsigned char> &b = ws ? *ws : op->ctx->mf_ws;
    const size_t bytes = sizeof(T) * (size_t)elems;
    if (b.n < bytes) b.alloc(bytes);
    return reinterpret_cast<T *>(b.p);
}

// LDS one block may use (160 KB on gfx950) and compute units (256), from the current device.
static int device_attr(hipDeviceAttribute_t a) {
    int dev = 0, v = 0;
    PE_HIP(hipGetDevice(&dev));
    PE_HIP(hipDeviceGetAttribute(&v, a, dev));
    return v;
}
static size_t lds_per_block() { return (size_t)device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock); }

template <int MT, typename T, bool SB, int VM, int PE, int ML>
static void launch_dprod_v(const MfView &M, const DprodArgs &A, const T *X, const T *V, const T *kv, T *W, size_t lds,
                           hipStream_t s) {
    auto kern = k_mf_dprod<MT, T, SB, VM, PE, ML>;
    // dynamic LDS above 64 KB, set once (a function-local static is initialised
    // once even when two host threads launch the kernel for the first time together)
    static const bool attr = [&] {
        PE_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds_per_block()));
        return true;
    }();
    (void)attr;
    // persistent grid: the blocks that are resident at once (registers and LDS)
    int per_cu = 0;
    PE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kDprodThreads, lds));
    static const i64 ncu = device_attr(hipDeviceAttributeMultiprocessorCount);
    const i64 ntiles = ((i64)A.ncol + 63) / 64;
    const unsigned grid = (unsigned)std::max<i64>(1, std::min<i64>(ntiles, ncu * std::max(1, per_cu)));
    kern<<<grid, kDprodThreads, lds, s>>>(A, M.tab, X, V, kv, W);
    PE_LAUNCH_CHECK();
}

// LDS plan of the fused pass for a pair dimension (gfx950: 160 KB per block): V in
// LDS beside two buffers where that fits (row stride 80, else 64), else beside one
// buffer, else V through the caches beside two buffers.
template <int MT, typename T>
struct DprodPlan {
    static constexpr size_t KP = MT * 16, BR = dprod_vc(MT) + 1, cap = 160 * 1024;
    static constexpr size_t vfull = KP * (dprod_vc(MT) + 1) * sizeof(T) + KP * sizeof(T);
    static constexpr size_t bufs(int n, int ld) { return (size_t)n * BR * ld * sizeof(double); }
    static constexpr int plan = vfull + bufs(2, 80) <= cap ? 0 : vfull + bufs(2, 64) <= cap ? 1
                              : vfull + bufs(1, 80) <= cap ? 2 : vfull + bufs(1, 64) <= cap ? 3 : 4;
    static constexpr bool SB = plan == 2 || plan == 3;
    static constexpr int VM = plan == 4 ? VM_GLOBAL : VM_FULL;
    static constexpr int ld = (plan == 1 || plan == 3) ? 64 : 80;
    static constexpr size_t lds = plan == 4 ? KP * sizeof(T) + bufs(2, 80) : vfull + bufs(SB ? 1 : 2, ld);
    static_assert(lds <= cap, "no LDS plan fits");
};

// How one call's vectors are laid out: C.nvec vectors (with C.dyn: at most that
// many, the count and offsets read on the device), row-major X[J * vstride + v0 + v]
// or column-major X[v * nd + J].
struct MfCall {
    i64 nd;
    int nvec, vstride, v0, colmajor;
    const long long *dyn;
};

// ml > 0: the live-row layout (row-major vectors, parity-sector scatter) with nbp
// columns per alpha string
template <int MT, typename T>
static void launch_dprod(const MfView &M, const MfCall &C, const T *X, const T *V, const T *kv, T *W, hipStream_t s,
                         int ml, i64 nbp) {
    typedef DprodPlan<MT, T> P;
    static const size_t cap = lds_per_block();
    const int ntot = M.nea + M.neb;
    PE_REQUIRE(ntot <= 128, PYCI_EUNSUPPORTED, "too many single excitations per determinant");
    PE_REQUIRE(C.nd * C.nvec < ((i64)1 << 31) - ((i64)1 << 20) && C.nd * C.vstride < ((i64)1 << 31),
               PYCI_EUNSUPPORTED, "too many columns for the matrix-free operator");
    PE_REQUIRE(!fixed_npair(MT) || M.npair == fixed_npair(MT), PYCI_EUNSUPPORTED,
               "pair count does not match the pair dimension");
    PE_REQUIRE(P::lds <= cap, PYCI_EUNSUPPORTED, "the device has too little LDS for the matrix-free product");
    // the folded k term needs a zero row of D below the k extent's end
    PE_REQUIRE(P::VM != VM_FULL || fixed_npair(MT) || M.npair < dprod_vc(MT), PYCI_EUNSUPPORTED,
               "pair count fills the pair dimension");
    DprodArgs A{(int)(C.nd * C.nvec), (int)M.nbs, C.nvec, M.npair, M.nea, ntot, C.vstride, C.v0, C.colmajor,
                (int)C.nd, C.dyn, (int)(M.na * M.nea), (int)(M.na * M.nea + M.nbs * M.neb), P::ld,
                (int)nbp, M.nea, M.live};
    if (ml > 0) {
        // the live rows of a string (nea of them) and the table entries per column
        // (2 nea) fix ml and PE together: ml = 2, 3, 4 with PE = 8, 12, 16
        PE_REQUIRE(!C.colmajor && !C.dyn && M.live && nbp % 64 == 0 && nbp >= M.nbs * C.nvec &&
                       M.na * nbp < ((i64)1 << 31) - ((i64)1 << 20) && ntot == 2 * M.nea &&
                       ml == (ntot <= 64 ? 2 : ntot <= 96 ? 3 : 4),
                   PYCI_EUNSUPPORTED, "live-row layout does not fit this call");
        A.ncol = (int)(M.na * nbp);
        if (ml == 2) {
            if constexpr (2 < pair_tiles(MT)) return launch_dprod_v<MT, T, P::SB, P::VM, 8, 2>(M, A, X, V, kv, W, P::lds, s);
        } else if (ml == 3) {
            if constexpr (3 < pair_tiles(MT)) return launch_dprod_v<MT, T, P::SB, P::VM, 12, 3>(M, A, X, V, kv, W, P::lds, s);
        } else if (ml == 4) {
            if constexpr (4 < pair_tiles(MT)) return launch_dprod_v<MT, T, P::SB, P::VM, 16, 4>(M, A, X, V, kv, W, P::lds, s);
        }
        throw EngineError(PYCI_EUNSUPPORTED, "no live-row product for this pair dimension");
    }
    if (ntot <= 64) launch_dprod_v<MT, T, P::SB, P::VM, 8, 0>(M, A, X, V, kv, W, P::lds, s);
    else if (ntot <= 96) launch_dprod_v<MT, T, P::SB, P::VM, 12, 0>(M, A, X, V, kv, W, P::lds, s);
    else launch_dprod_v<MT, T, P::SB, P::VM, 16, 0>(M, A, X, V, kv, W, P::lds, s);
}

// ev (optional): 3 events recorded before, between and after the two passes.
template <int NV, typename T>
static void launch_passes(const MfView &M, const MfCall &C, const T *X, T *W, T *Y, const T *V, const T *kv,
                          hipStream_t s, hipEvent_t *ev, int sector) {
    if (ev) PE_HIP(hipEventRecord(ev[0], s));
    const bool one_vec = C.nvec == 1 && C.vstride == 1 && C.v0 == 0 && !C.colmajor && !C.dyn;
    // the live-row layout for row-major vectors in a sector: NB nvec columns per alpha
    // string, padded to a multiple of 64
    const int ml = (sector != 0 && !C.colmajor && !C.dyn) ? M.ml : 0;
    const i64 nbp = ml ? (M.nbs * C.nvec + 63) / 64 * 64 : 0;
    switch (M.kpad / 16) {
        case 1: launch_dprod<1, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 2: launch_dprod<2, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 3: launch_dprod<3, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 4: launch_dprod<4, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 5: launch_dprod<5, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 6: launch_dprod<6, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 7: launch_dprod<7, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        case 8: launch_dprod<8, T>(M, C, X, V, kv, W, s, ml, nbp); break;
        default: throw EngineError(PYCI_EUNSUPPORTED, "pair dimension too large for the matrix-free operator");
    }
    if (ev) PE_HIP(hipEventRecord(ev[1], s));
    const i64 nd = C.nd;
    if (sector != 0) {
        // alpha terms (every block of 256 consecutive determinants shares few alpha
        // strings, so their table rows are common and the E reads contiguous), then
        // the parity symmetrisation
        if (one_vec) {
            const int nchunk = (int)((M.nbs + 63) / 64);
            const i64 waves = M.na * nchunk;
            if (ml > 0)
                k_mf_alpha<T><<<(unsigned)((waves + 3) / 4), 256, 0, s>>>((int)M.na, (int)M.nbs, (int)nbp, nchunk,
                                                                         M.nea, 16 * ml, M.tab_al, W, Y);
            else
                k_mf_alpha<T><<<(unsigned)((waves + 3) / 4), 256, 0, s>>>((int)M.na, (int)M.nbs, (int)M.nbs, nchunk,
                                                                         M.nea, M.kpad, M.tab_a, W, Y);
        } else {
            k_mf_sigma<NV, T, false><<<grid_for(nd, 256), 256, 0, s>>>(
                nd, M.nbs, M.nea, M.neb, ml ? 16 * ml : M.kpad, ml ? M.tab_al : M.tab_a, M.tab_b, C.nvec, C.vstride,
                C.v0, 1, C.colmajor, C.dyn, W, Y, nbp);
        }
        PE_LAUNCH_CHECK();
        if (C.nvec == 1 && C.vstride == 1 && C.v0 == 0 && !C.colmajor && !C.dyn) {
            const i64 nt = (M.nbs + 31) / 32;
            k_mf_symm_tiles<T><<<(unsigned)(nt * (nt + 1) / 2), 256, 0, s>>>(Y, (int)M.nbs, (T)sector);
        } else {
            k_mf_symm<T><<<grid_for(nd, 256), 256, 0, s>>>(Y, M.nbs, C.nvec, C.vstride, C.v0, C.colmajor, C.dyn,
                                                           (T)sector);
        }
        PE_LAUNCH_CHECK();
        if (ev) PE_HIP(hipEventRecord(ev[2], s));
        return;
    }
    const bool tiled = (M.nbs + 15) / 16 >= 16;  // enough column tiles to spread over the XCDs
    const unsigned g = tiled ? tiled_grid(M.na, M.nbs) : grid_for(nd, 256);
    if (tiled)
        k_mf_sigma<NV, T, true><<<g, 256, 0, s>>>(nd, M.nbs, M.nea, M.neb, M.kpad, M.tab_a, M.tab_b, C.nvec,
                                                  C.vstride, C.v0, 0, C.colmajor, C.dyn, W, Y, 0);
    else
        k_mf_sigma<NV, T, false><<<g, 256, 0, s>>>(nd, M.nbs, M.nea, M.neb, M.kpad, M.tab_a, M.tab_b, C.nvec,
                                                   C.vstride, C.v0, 0, C.colmajor, C.dyn, W, Y, 0);
    PE_LAUNCH_CHECK();
    if (ev) PE_HIP(hipEventRecord(ev[2], s));
}

template <typename T>
static void launch_passes_nv(const MfView &M, const MfCall &C, const T *X, T *W, T *Y, const T *V, const T *kv,
                             hipStream_t s, hipEvent_t *ev, int sector) {
    if (C.nvec == 1) launch_passes<1, T>(M, C, X, W, Y, V, kv, s, ev, sector);
    else if (C.nvec == 2) launch_passes<2, T>(M, C, X, W, Y, V, kv, s, ev, sector);
    else if (C.nvec <= 4) launch_passes<4, T>(M, C, X, W, Y, V, kv, s, ev, sector);
    else if (C.nvec <= 8) launch_passes<8, T>(M, C, X, W, Y, V, kv, s, ev, sector);
    else launch_passes<16, T>(M, C, X, W, Y, V, kv, s, ev, sector);
}

static MfView view_of(const DirectOp *D) {
    MfView M{D->tab.p, D->tab_a.p, D->tab_b.p, D->na, D->nbs, D->npair, D->kpad, D->nea, D->neb};
    M.ml = D->ml;
    M.nbp = D->nbp;
    M.live = D->live.p;
    M.tab_al = D->tab_al.p;
    return M;
}

// Workspace values of one sigma of nc vectors: kpad rows per column, or for the
// live-row layout (row-major vectors in a parity sector) 16 ml rows per column of the
// padded layout (NB nc columns per alpha string rounded up to 64).
static i64 workspace_elems(const DirectOp *D, i64 nd, int nc, int sector) {
    i64 e = ((nd * nc + 63) / 64 * 64) * (i64)D->kpad;
    if (D->ml > 0 && sector != 0) e = std::max<i64>(e, D->na * ((D->nbs * nc + 63) / 64 * 64) * 16 * D->ml);
    return e;
}

// The vectors go through the two passes in chunks whose E workspace stays within
// 16 GB (one vector at least): at nbasis 14 one FP64 vector already needs 10.6 GB.
// sector = +1 / -1: every vector lies in that parity sector of the alpha <-> beta
// exchange (only where nup == ndn; the solver's sector runs pass it), which halves
// the scatter pass (k_mf_symm); 0: any vectors.
template <typename T>
void direct_sigma(const pyci_op *op, int nvec, const T *X, T *Y, hipStream_t s, hipEvent_t *ev, int sector,
                  DevBuf<unsigned char> *ws) {
    DirectOp *D = op->direct.get();
    if (D->nup != D->ndn) sector = 0;
    PE_REQUIRE(nvec >= 1 && nvec <= 16, PYCI_EUNSUPPORTED, "the matrix-free operator takes 1..16 vectors");
    const i64 nd = op->nrow;
    const i64 per_vec = ((nd + 63) / 64 * 64) * (i64)D->kpad * (i64)sizeof(T);
    const int chunk = (int)std::max<i64>(1, std::min<i64>(nvec, ((i64)16 << 30) / per_vec));
    T *W = direct_workspace<T>(op, workspace_elems(D, nd, chunk, sector), ws);
    const T *V = std::is_same<T, double>::value ? (const T *)D->V64.p : (const T *)D->V32.p;
    const T *kv = std::is_same<T, double>::value ? (const T *)D->k64.p : (const T *)D->k32.p;
    const MfView M = view_of(D);
    for (int v0 = 0; v0 < nvec; v0 += chunk) {
        const int nc = std::min(chunk, nvec - v0);
        hipEvent_t *e = (ev && v0 == 0 && nc == nvec) ? ev : nullptr;
        launch_passes_nv<T>(M, MfCall{nd, nc, nvec, v0, 0, nullptr}, X, W, Y, V, kv, s, e, sector);
    }
}

// Sigma of a batched-solver problem (batch.hip): up to nvec_max column-major
// vectors, their count and the offsets of X and Y read on the device from dyn at
// run time (a launch inside a captured graph that finds a count of 0 returns at
// once). W holds kpad x (nd nvec_max) values.
template <typename T>
void mf_sigma_batched(const MfView &M, i64 nd, int nvec_max, const T *X, T *Y, T *W, const T *V, const T *kv,
                      const long long *dyn, int sector, hipStream_t s) {
    launch_passes_nv<T>(M, MfCall{nd, nvec_max, 1, 0, 1, dyn}, X, W, Y, V, kv, s, nullptr, sector);
}

template void mf_sigma_batched<double>(const MfView &, i64, int, const double *, double *, double *, const double *,
                                       const double *, const long long *, int, hipStream_t);
template void mf_sigma_batched<float>(const MfView &, i64, int, const float *, float *, float *, const float *,
                                      const float *, const long long *, int, hipStream_t);

// ---------------------------------------------------------------------------
// Allocation-free sigma: the E workspace in a fixed device array instead of the
// context's grown-on-demand buffer, for callers that cannot allocate (code under
// graph capture, serving loops). The array holds kNoallocElems doubles (256 MiB)
// on every device the library runs on, for the life of the process; calls that
// overlap on one device share it and race. It needs no clearing: the scatter
// pass reads only what the product pass of the same call wrote.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocElems = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kNoallocElems];

// Device address of `scratch` on the current device: looked up once per device
// (the name is only a host-side handle; each device holds its own copy).
double *noalloc_scratch(int device) {
    static std::mutex mu;
    static std::vector<double *> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)device >= ptrs.size()) ptrs.resize((size_t)device + 1, nullptr);
    if (!ptrs[(size_t)device]) {
        double *p = nullptr;
        PE_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        ptrs[(size_t)device] = p;
    }
    return ptrs[(size_t)device];
}

size_t noalloc_scratch_bytes() { return kNoallocElems * sizeof(double); }

void direct_sigma_noalloc(const pyci_op *op, int nvec, const double *X, double *Y, double *ws, hipStream_t s) {
    const DirectOp *D = op->direct.get();
    PE_REQUIRE(nvec >= 1 && nvec <= 16, PYCI_EUNSUPPORTED, "the matrix-free operator takes 1..16 vectors");
    const i64 nd = op->nrow;
    const i64 need = ((nd * nvec + 63) / 64 * 64) * (i64)D->kpad;
    PE_REQUIRE((size_t)need <= kNoallocElems, PYCI_EUNSUPPORTED,
               "the allocation-free product needs " + std::to_string(need * 8) + " bytes of workspace for " +
                   std::to_string(nvec) + " vector(s); its fixed workspace holds " +
                   std::to_string(kNoallocElems * 8) + " bytes");
    launch_passes_nv<double>(view_of(D), MfCall{nd, nvec, nvec, 0, 0, nullptr}, X, ws, Y, D->V64.p, D->k64.p, s,
                             nullptr, 0);
}

template void direct_sigma<double>(const pyci_op *, int, const double *, double *, hipStream_t, hipEvent_t *, int,
                                   DevBuf<unsigned char> *);
template void direct_sigma<float>(const pyci_op *, int, const float *, float *, hipStream_t, hipEvent_t *, int,
                                  DevBuf<unsigned char> *);

// Per-pass timing of nrep matrix-free sigmas on random X (see pyci_op_bench_sigma_parts).
// One vector with nup == ndn is timed as the solver runs it, in a parity sector (the
// sector's sigma of random X costs what it costs on a vector of the sector).
// Algorithmic bytes and flops per launch, with N = nd * nvec columns and T the value size:
//   gather + product: reads X (T N), writes E (T R N); 2 R npair N flops, R = npair, or
//                     in a sector with the live-row layout nea (the rows stored)
//   scatter pass:     reads E (T R N), writes Y (T N); in a sector it then reads and
//                     writes Y once more (the parity symmetrisation)
void direct_bench_parts(pyci_op *op, i64 nvec, i64 nrep, int fp32, float *ms, double *bytes, double *flops) {
    DirectOp *D = op->direct.get();
    hipStream_t s = op->ctx->stream;
    const i64 nd = op->nrow;
    // one chunk of vectors (direct_sigma's 16 GB workspace budget), so that the
    // events bracket exactly the two passes
    const i64 per_vec = ((nd + 63) / 64 * 64) * (i64)D->kpad * (fp32 ? 4 : 8);
    nvec = std::max<i64>(1, std::min<i64>(nvec, ((i64)16 << 30) / per_vec));
    const i64 N = nd * nvec;
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
    std::vector<hipEvent_t> ev((size_t)(3 * nrep));
    for (auto &e : ev) PE_HIP(hipEventCreate(&e));
    const int sector = (D->nup == D->ndn && nvec == 1) ? 1 : 0;
    auto run = [&](hipEvent_t *e) {
        if (fp32) direct_sigma<float>(op, (int)nvec, x32.p, y32.p, s, e, sector);
        else direct_sigma<double>(op, (int)nvec, x.p, y.p, s, e, sector);
    };
    try {
        run(nullptr);  // warm-up (workspace allocation)
        for (i64 r = 0; r < nrep; ++r) run(&ev[(size_t)(3 * r)]);
        PE_HIP(hipEventSynchronize(ev.back()));
        double acc[2] = {0, 0};
        for (i64 r = 0; r < nrep; ++r)
            for (int p = 0; p < 2; ++p) {
                float t = 0.f;
                PE_HIP(hipEventElapsedTime(&t, ev[(size_t)(3 * r + p)], ev[(size_t)(3 * r + p + 1)]));
                acc[p] += t;
            }
        const double T = fp32 ? 4.0 : 8.0, P = (double)D->npair, NN = (double)N;
        const double R = (sector && D->ml) ? (double)D->nea : P;
        for (int p = 0; p < 2; ++p) ms[p] = (float)(acc[p] / (double)nrep);
        bytes[0] = T * NN * (1.0 + R);
        bytes[1] = T * NN * (R + 1.0 + (sector ? 2.0 : 0.0));
        flops[0] = 2.0 * R * P * NN + 2.0 * (double)(D->nea + D->neb) * NN;
        flops[1] = 2.0 * (double)(sector ? D->nea : D->nea + D->neb) * NN;
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
