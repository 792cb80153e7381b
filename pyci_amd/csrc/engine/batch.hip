// Batched eigensolves of many uneven FullCI problems (full spaces, colex order).
//
// No counterpart in the reference, whose solve_ci (sparseop.cpp:114-146) runs one
// problem per call; each problem here is the reference's SparseOp(ham, wfn) +
// solve_ci(n) on the full FullCI space of its own integrals.
//
// All problems advance together, one Davidson iteration per step, and every
// step is ONE kernel launch over all problems: a block map sends each
// workgroup to (problem, row range) and kernels skip problems that converged.
// All control (convergence, thick restart, FP32 -> FP64 switch, dependent
// vectors) is decided on the device from per-problem state in HBM, so an
// iteration is a fixed launch sequence that is captured once into a hipGraph
// and replayed; the host only polls the number of active problems.
//
// Per problem: basis V and images W (mmax x nd, one vector contiguous), Ritz
// vectors X, WX (2k x nd), corrections T (k x nd), the Rayleigh matrix G, and the
// matrix-free sigma of direct.hip (string tables, pair integrals, D/E workspace).
// Iteration: Rayleigh-Ritz (Jacobi in LDS, one workgroup per problem) -> Ritz
// vectors, residuals, preconditioned corrections -> control -> restart/copy ->
// CGS2 against the basis + Cholesky-QR of the new block -> sigma of the new
// columns (FP32 or FP64 per problem) -> new Rayleigh-matrix entries.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace pe {

constexpr int BK_MAX = 8;   // roots per problem
constexpr int BM_MAX = 48;  // subspace size
constexpr int NB2 = 16;     // new columns at once (the start block holds up to 2k)
constexpr int ST = 16;      // ints of per-problem state
enum { S_M = 0, S_C0, S_NADD, S_KK, S_ACTIVE, S_PHASE, S_RESTART, S_ITERS, S_REINIT, S_FAIL, S_STALL, S_NSIG,
       S_DONE };

struct BProb {
    i64 nd, NB;
    int n, npair, kpad, nea, neb, mt;
    i64 tab_a, tab_b;  // offsets into the table pool (int2)
    i64 vint, kv;      // offsets into the pair-integral pools
    i64 vec;           // offset of nd-long arrays (diagonal)
    i64 basis;         // offset of V and W (mmax * nd each)
    i64 ritz;          // offset of X and WX (2k * nd each)
    i64 corr;          // offset of T (k * nd)
    i64 s32;           // offset of the FP32 scratch (kb0 * nd each)
    i64 work, ldw;     // D/E workspace offset, kpad rows x ldw columns
    i64 rb0, nrb;      // row blocks [rb0, rb0 + nrb)
    double ecore;
    i64 swap_nb;       // sector != 0: beta-string count of the alpha <-> beta exchange P
    int sector;        // 0, or +1 / -1: this sub-problem's parity sector of P
    int orig;          // the caller's problem index
    int big;           // sigma by the fused pass of direct.hip (own launches) instead of k_bmf_*
};

__device__ inline i64 swap_partner(i64 J, i64 nb) { return (J % nb) * nb + J / nb; }

// Device view of the whole batch.
struct BView {
    int nprob, k, kb0, mmax, S, RB;
    double tol, switch_tol;
    int maxiter;
    const BProb *prob;
    const int *row_blk;      // row block -> problem
    const i64 *row_first;    // row block -> first row
    const int *col_blk;      // sigma column block -> problem (256 columns)
    const i64 *col_first;
    const int *tile_blk;     // product tile -> problem (128 columns)
    const i64 *tile_first;
    const int2 *tab;
    const double *V64p;
    const double *k64p;
    const float *V32p;
    const float *k32p;
    const double *diag;
    double *V, *W, *X, *WX, *T;
    float *S32x, *S32y;
    double *work64;
    float *work32;
    int *st;                 // nprob x ST
    int *sel;                // nprob x BK_MAX
    double *theta, *Y, *G, *res, *best, *out, *C, *Rinv, *nrm0, *energy;
    double *partial;         // row blocks x S
    unsigned long long *tconv;
    int *nactive;
};

// ---------------------------------------------------------------------------
// Rayleigh-Ritz: cyclic Jacobi of G (m x m) in LDS, one 64-thread workgroup per
// problem. Decides the restart (m + k > mmax) and the number of Ritz vectors kk.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_bjac(BView b) {
    const int p = blockIdx.x, t = threadIdx.x;
    int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const int m = st[S_M], k = b.k, mmax = b.mmax;
    __shared__ double a[BM_MAX][BM_MAX + 1], v[BM_MAX][BM_MAX + 1];
    __shared__ int order[BM_MAX];
    const double *G = b.G + (i64)p * BM_MAX * BM_MAX;
    for (int idx = t; idx < m * m; idx += 64) {
        const int i = idx / m, j = idx % m;
        a[i][j] = G[i * BM_MAX + j];
        v[i][j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = t; idx < m * m; idx += 64) {
            const int i = idx / m, j = idx % m;
            const double x = a[i][j] * a[i][j];
            tot += x;
            if (i != j) off += x;
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            tot += __shfl_xor(tot, o);
        }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int pp = 0; pp < m - 1; ++pp)
            for (int qq = pp + 1; qq < m; ++qq) {
                const double apq = a[pp][qq];
                if (apq == 0.0) continue;
                const double app = a[pp][pp], aqq = a[qq][qq];
                const double tau = (aqq - app) / (2.0 * apq);
                const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
                __syncthreads();
                if (t < m) {
                    const double arp = a[t][pp], arq = a[t][qq];
                    a[t][pp] = c * arp - s * arq;
                    a[t][qq] = s * arp + c * arq;
                }
                __syncthreads();
                if (t < m) {
                    const double apr = a[pp][t], aqr = a[qq][t];
                    a[pp][t] = c * apr - s * aqr;
                    a[qq][t] = s * apr + c * aqr;
                    const double vrp = v[t][pp], vrq = v[t][qq];
                    v[t][pp] = c * vrp - s * vrq;
                    v[t][qq] = s * vrp + c * vrq;
                }
                __syncthreads();
            }
    }
    if (t == 0) {
        for (int i = 0; i < m; ++i) order[i] = i;
        for (int i = 1; i < m; ++i) {  // stable insertion sort, ascending
            const int x = order[i];
            int j = i - 1;
            while (j >= 0 && a[order[j]][order[j]] > a[x][x]) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = x;
        }
        const int restart = (m + k > mmax) ? 1 : 0;
        int kk = k;
        if (restart) kk = max(k, min(m, min(2 * k, mmax - k)));
        st[S_RESTART] = restart;
        st[S_KK] = kk;
    }
    __syncthreads();
    const int kk = st[S_KK];
    double *th = b.theta + (i64)p * BM_MAX, *Y = b.Y + (i64)p * BM_MAX * 2 * BK_MAX;
    for (int j = t; j < m; j += 64) th[j] = a[order[j]][order[j]];
    for (int idx = t; idx < m * kk; idx += 64) {
        const int i = idx / kk, j = idx % kk;
        Y[i * 2 * BK_MAX + j] = v[i][order[j]];
    }
}

// ---------------------------------------------------------------------------
// Ritz vectors X = V Y, WX = W Y (kk of them), corrections T = (WX - theta X) /
// (theta - diag) and residual-norm partials for the k roots. Row-mapped.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_britz(BView b) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int m = st[S_M], kk = st[S_KK], k = b.k;
    const double *th = b.theta + (i64)p * BM_MAX, *Y = b.Y + (i64)p * BM_MAX * 2 * BK_MAX;
    const double *V = b.V + P.basis, *W = b.W + P.basis;
    double *X = b.X + P.ritz, *WX = b.WX + P.ritz, *T = b.T + P.corr;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    double ss[BK_MAX];
#pragma unroll
    for (int j = 0; j < BK_MAX; ++j) ss[j] = 0.0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256) {
        double x[2 * BK_MAX], w[2 * BK_MAX];
#pragma unroll
        for (int j = 0; j < 2 * BK_MAX; ++j) x[j] = w[j] = 0.0;
        for (int i = 0; i < m; ++i) {
            const double vv = V[(i64)i * P.nd + r], ww = W[(i64)i * P.nd + r];
#pragma unroll
            for (int j = 0; j < 2 * BK_MAX; ++j)
                if (j < kk) {
                    const double y = Y[i * 2 * BK_MAX + j];
                    x[j] += vv * y;
                    w[j] += ww * y;
                }
        }
        const double d = b.diag[P.vec + r];
#pragma unroll
        for (int j = 0; j < 2 * BK_MAX; ++j)
            if (j < kk) {
                X[(i64)j * P.nd + r] = x[j];
                WX[(i64)j * P.nd + r] = w[j];
            }
#pragma unroll
        for (int j = 0; j < BK_MAX; ++j)
            if (j < k) {
                const double res = w[j] - th[j] * x[j];
                double den = th[j] - d;
                if (fabs(den) < 1e-8) den = (den < 0.0) ? -1e-8 : 1e-8;
                T[(i64)j * P.nd + r] = res / den;
                ss[j] += res * res;
            }
    }
    __shared__ double red[4][BK_MAX];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < BK_MAX; ++j) {
        double v = ss[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)k)
        b.partial[(i64)blockIdx.x * b.S + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// ---------------------------------------------------------------------------
// Control, one workgroup per problem: residual norms, convergence, phase
// switch, stagnation guard, restart bookkeeping, which roots get corrections.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_bctl(BView b) {
    const int p = blockIdx.x;
    int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int k = b.k;
    __shared__ double rs[BK_MAX];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = wave; j < k; j += 4) {
        double v = 0.0;
        for (i64 q = lane; q < P.nrb; q += 64) v += b.partial[(P.rb0 + q) * b.S + j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) rs[j] = sqrt(v);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double eps23 = 3.666852862501036e-11;  // (2^-52)^(2/3)
    const double *th = b.theta + (i64)p * BM_MAX;
    double *res = b.res + (i64)p * BK_MAX, *best = b.best + (i64)p * BK_MAX;
    int *sel = b.sel + p * BK_MAX;
    int nun = 0;
    bool floor = true, improved = false, near_floor = true;
    for (int j = 0; j < k; ++j) {
        res[j] = rs[j];
        const double lim = b.tol * fmax(fabs(th[j]), eps23);
        if (!(rs[j] <= lim)) sel[nun++] = j;
        if (!(rs[j] <= fmax(lim, b.switch_tol * fmax(fabs(th[j]), 1.0)))) floor = false;
        if (!(rs[j] <= 1e-4 * fmax(fabs(th[j]), 1.0))) near_floor = false;
        if (rs[j] < 0.9 * best[j]) {
            best[j] = rs[j];
            improved = true;
        }
    }
    st[S_STALL] = improved ? 0 : st[S_STALL] + 1;
    st[S_ITERS] += 1;
    st[S_REINIT] = 0;
    const bool fp32 = st[S_PHASE] == 1;
    bool done = (nun == 0);
    // as the single-problem solver: a stall hands the FP32 phase over only near its
    // rounding floor (every residual within 1e-4 |theta|), or after 40 flat iterations
    bool switch_phase = fp32 && (done || floor || (st[S_STALL] >= 8 && near_floor) || st[S_STALL] >= 40);
    if (!fp32 && !done && st[S_STALL] >= 25) {
        // as the single-problem solver: accept a residual at the rounding floor, fail
        // only after 300 flat iterations
        bool tiny = true;
        for (int j = 0; j < k; ++j)
            if (!(rs[j] <= 1e-7 * fmax(fabs(th[j]), 1.0))) tiny = false;
        if (tiny) done = true;
        else if (st[S_STALL] >= 300) {
            st[S_FAIL] = 1;
            done = true;
        }
    }
    if (!done && st[S_ITERS] >= b.maxiter) {
        st[S_FAIL] = 1;
        done = true;
        switch_phase = false;
    }
    if (switch_phase) {
        // FP64 residual-correction phase: restart on the FP32 phase's Ritz vectors
        st[S_PHASE] = 0;
        st[S_REINIT] = 1;
        st[S_RESTART] = 0;
        st[S_M] = 0;
        st[S_C0] = 0;
        st[S_NADD] = k;
        st[S_STALL] = 0;
        for (int j = 0; j < k; ++j) best[j] = 1e300;
        return;
    }
    if (done) {
        st[S_ACTIVE] = 0;
        st[S_NADD] = 0;
        st[S_RESTART] = 0;
        for (int j = 0; j < k; ++j) b.energy[(i64)p * BK_MAX + j] = th[j] + P.ecore;
        b.tconv[p] = wall_clock64();
        return;
    }
    int m = st[S_M];
    if (st[S_RESTART]) {
        const int kk = st[S_KK];
        double *G = b.G + (i64)p * BM_MAX * BM_MAX;
        for (int i = 0; i < kk; ++i)
            for (int j = 0; j < kk; ++j) G[i * BM_MAX + j] = (i == j) ? th[i] : 0.0;
        m = kk;
        st[S_M] = m;
    }
    st[S_C0] = m;
    st[S_NADD] = min(nun, b.mmax - m);
}

// Restart copies and the new columns. Row-mapped.
__global__ void __launch_bounds__(256) k_bmove(BView b) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int restart = st[S_RESTART], kk = st[S_KK], c0 = st[S_C0], nadd = st[S_NADD], reinit = st[S_REINIT];
    const int *sel = b.sel + p * BK_MAX;
    double *V = b.V + P.basis, *W = b.W + P.basis;
    const double *X = b.X + P.ritz, *WX = b.WX + P.ritz, *T = b.T + P.corr;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256) {
        if (restart && !reinit)
            for (int j = 0; j < kk; ++j) {
                V[(i64)j * P.nd + r] = X[(i64)j * P.nd + r];
                W[(i64)j * P.nd + r] = WX[(i64)j * P.nd + r];
            }
        for (int a = 0; a < nadd; ++a) {
            double x;
            if (reinit) x = X[(i64)a * P.nd + r];
            else if (P.sector != 0)  // corrections stay in the sector
                x = 0.5 * (T[(i64)sel[a] * P.nd + r] + P.sector * T[(i64)sel[a] * P.nd + swap_partner(r, P.swap_nb)]);
            else x = T[(i64)sel[a] * P.nd + r];
            V[(i64)(c0 + a) * P.nd + r] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// Tall-skinny products. mode 0: V[:, 0:c0]^T N; 1: N^T N; 2: V[:, 0:c0+nadd]^T W[:, c0:c0+nadd]
// with N = V[:, c0:c0+nadd]. Partials per row block, 8 x 8 tile per blockIdx.y.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_bgram(BView b, int mode) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int c0 = st[S_C0], nadd = st[S_NADD];
    const int ma = (mode == 0) ? c0 : (mode == 1 ? nadd : c0 + nadd), mb = nadd;
    if (ma == 0 || mb == 0) return;
    const int tj = (mb + 7) / 8;
    const int i0 = (blockIdx.y / tj) * 8, j0 = (blockIdx.y % tj) * 8;
    if (i0 >= ma || j0 >= mb) return;
    const int ni = min(8, ma - i0), nj = min(8, mb - j0);
    const double *A = b.V + P.basis + (mode == 1 ? (i64)c0 * P.nd : 0);
    const double *B = (mode == 2 ? b.W : b.V) + P.basis + (i64)c0 * P.nd;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256) {
        double x[8], y[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = (i < ni) ? A[(i64)(i0 + i) * P.nd + r] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) y[j] = (j < nj) ? B[(i64)(j0 + j) * P.nd + r] : 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] += x[i] * y[j];
    }
    __shared__ double red[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            double v = acc[i][j];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) red[wave][i * 8 + j] = v;
        }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int i = threadIdx.x / 8, j = threadIdx.x % 8;
        if (i < ni && j < nj)
            b.partial[(i64)blockIdx.x * b.S + (i0 + i) * mb + j0 + j] =
                ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
    }
}

// out[p][s] = sum over the problem's row blocks of partial[s]; a wave per slot.
__global__ void __launch_bounds__(256) k_bred(BView b, int mode) {
    const int p = blockIdx.x;
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int c0 = st[S_C0], nadd = st[S_NADD];
    const int ma = (mode == 0) ? c0 : (mode == 1 ? nadd : c0 + nadd);
    const int cnt = ma * nadd;
    const int s = blockIdx.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= cnt) return;
    double v = 0.0;
    for (i64 q = lane; q < P.nrb; q += 64) v += b.partial[(P.rb0 + q) * b.S + s];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) b.out[(i64)p * b.S + s] = v;
}

// N -= V[:, 0:c0] C (C = out, c0 x nadd). Row-mapped.
__global__ void __launch_bounds__(256) k_bupd(BView b) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int c0 = st[S_C0], nadd = st[S_NADD];
    if (c0 == 0 || nadd == 0) return;
    const double *C = b.out + (i64)p * b.S;
    double *V = b.V + P.basis;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256)
        for (int a = 0; a < nadd; ++a) {
            double acc = 0.0;
            for (int i = 0; i < c0; ++i) acc += V[(i64)i * P.nd + r] * C[i * nadd + a];
            V[(i64)(c0 + a) * P.nd + r] -= acc;
        }
}

// Per problem (one thread): save = 1 keeps the norms of the new columns before
// orthogonalisation; save = 0 forms R^-1 of the Cholesky factor of N^T N, keeping
// only columns whose norm did not collapse below 1e-10 of the saved one.
__global__ void k_bchol(BView b, int save) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.nprob) return;
    int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const int nadd = st[S_NADD];
    const double *M = b.out + (i64)p * b.S;
    double *n0 = b.nrm0 + (i64)p * NB2;
    if (save) {
        for (int a = 0; a < nadd; ++a) n0[a] = sqrt(fmax(M[a * nadd + a], 0.0));
        return;
    }
    double R[NB2][NB2];
    int keep[NB2], nk = 0;
    for (int a = 0; a < nadd; ++a) {
        // column a against the kept ones: r_ba = (M_ba - sum_c r_cb r_ca) / r_bb
        double col[NB2];
        for (int q = 0; q < nk; ++q) {
            const int bq = keep[q];
            double v = M[bq * nadd + a];
            for (int c = 0; c < q; ++c) v -= R[c][q] * col[c];
            col[q] = v / R[q][q];
        }
        double d2 = M[a * nadd + a];
        for (int q = 0; q < nk; ++q) d2 -= col[q] * col[q];
        const double d = sqrt(fmax(d2, 0.0));
        if (d > 1e-10 * fmax(n0[a], 1e-300) && d > 1e-300) {
            for (int q = 0; q < nk; ++q) R[q][nk] = col[q];
            R[nk][nk] = d;
            keep[nk++] = a;
        }
    }
    // R^-1 (nk x nk upper triangular) as coefficients over the original columns
    double Ri[NB2][NB2];
    for (int j = 0; j < nk; ++j) {
        for (int i = 0; i < nk; ++i) Ri[i][j] = 0.0;
        Ri[j][j] = 1.0 / R[j][j];
        for (int i = j - 1; i >= 0; --i) {
            double v = 0.0;
            for (int q = i + 1; q <= j; ++q) v += R[i][q] * Ri[q][j];
            Ri[i][j] = -v / R[i][i];
        }
    }
    double *Rinv = b.Rinv + (i64)p * NB2 * NB2;  // [original column][kept column]
    for (int a = 0; a < NB2; ++a)
        for (int j = 0; j < NB2; ++j) Rinv[a * NB2 + j] = 0.0;
    for (int i = 0; i < nk; ++i)
        for (int j = 0; j < nk; ++j) Rinv[keep[i] * NB2 + j] = Ri[i][j];
    st[S_NADD] = nk;
    st[S_DONE] = nadd;  // columns present before the QR (read by k_bapplyR)
}

// N <- N R^-1 over the kept columns, row by row in place. Row-mapped.
__global__ void __launch_bounds__(256) k_bapplyR(BView b) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const BProb P = b.prob[p];
    const int c0 = st[S_C0], nk = st[S_NADD], nold = st[S_DONE];
    if (nold == 0) return;
    const double *Rinv = b.Rinv + (i64)p * NB2 * NB2;
    double *V = b.V + P.basis + (i64)c0 * P.nd;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256) {
        double x[NB2], y[NB2];
        for (int a = 0; a < NB2; ++a) x[a] = (a < nold) ? V[(i64)a * P.nd + r] : 0.0;
        for (int j = 0; j < NB2; ++j) {
            double v = 0.0;
            for (int a = 0; a < NB2; ++a)
                if (a < nold) v += x[a] * Rinv[a * NB2 + j];
            y[j] = v;
        }
        for (int j = 0; j < NB2; ++j)
            if (j < nk) V[(i64)j * P.nd + r] = y[j];
    }
}

// New Rayleigh-matrix entries (out: (c0+nadd) x nadd) and the basis size; counts
// the sigma products. One thread per problem.
__global__ void k_bgupd(BView b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.nprob) return;
    int *st = b.st + p * ST;
    if (!st[S_ACTIVE]) return;
    const int c0 = st[S_C0], nadd = st[S_NADD];
    const double *h = b.out + (i64)p * b.S;
    double *G = b.G + (i64)p * BM_MAX * BM_MAX;
    for (int i = 0; i < c0 + nadd; ++i)
        for (int a = 0; a < nadd; ++a) {
            const double x = h[i * nadd + a];
            G[i * BM_MAX + c0 + a] = x;
            G[(c0 + a) * BM_MAX + i] = x;
        }
    for (int a = 0; a < nadd; ++a)
        for (int c = a + 1; c < nadd; ++c) {
            const double x = 0.5 * (G[(c0 + a) * BM_MAX + c0 + c] + G[(c0 + c) * BM_MAX + c0 + a]);
            G[(c0 + a) * BM_MAX + c0 + c] = G[(c0 + c) * BM_MAX + c0 + a] = x;
        }
    st[S_M] = c0 + nadd;
    st[S_NSIG] += nadd;
}

__global__ void k_bcount(BView b) {
    __shared__ int c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    int mine = 0;
    for (int p = threadIdx.x; p < b.nprob; p += blockDim.x) mine += b.st[p * ST + S_ACTIVE];
    atomicAdd(&c, mine);
    __syncthreads();
    if (threadIdx.x == 0) *b.nactive = c;
}

// ---------------------------------------------------------------------------
// Batched matrix-free sigma of the new columns (see direct.hip for the passes).
// Columns c = v * nd + J of the problem's new block; FP32 problems go through
// FP32 copies of the columns.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline bool sigma_here(const int *st) {
    return st[S_ACTIVE] && st[S_NADD] > 0 && (st[S_PHASE] == (sizeof(T) == 4 ? 1 : 0));
}

template <typename T>
__device__ inline const T *xcols(const BView &b, const BProb &P, int c0) {
    if constexpr (sizeof(T) == 8) return b.V + P.basis + (i64)c0 * P.nd;
    else return b.S32x + P.s32;
}
template <typename T>
__device__ inline T *ycols(const BView &b, const BProb &P, int c0) {
    if constexpr (sizeof(T) == 8) return b.W + P.basis + (i64)c0 * P.nd;
    else return b.S32y + P.s32;
}
template <typename T>
__device__ inline T *wspace(const BView &b, const BProb &P) {
    if constexpr (sizeof(T) == 8) return b.work64 + P.work;
    else return b.work32 + P.work;
}

// to = 0: FP64 new columns -> FP32 scratch; to = 1: FP32 images -> W. Row-mapped.
__global__ void __launch_bounds__(256) k_bcvt(BView b, int to) {
    const int p = b.row_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!sigma_here<float>(st)) return;
    const BProb P = b.prob[p];
    const int c0 = st[S_C0], nadd = st[S_NADD];
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256)
        for (int a = 0; a < nadd; ++a) {
            const i64 o = (i64)a * P.nd + r;
            if (to == 0) b.S32x[P.s32 + o] = (float)b.V[P.basis + (i64)c0 * P.nd + o];
            else b.W[P.basis + (i64)c0 * P.nd + o] = (double)b.S32y[P.s32 + o];
        }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bmf_d(BView b) {
    const int p = b.col_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!sigma_here<T>(st)) return;
    const BProb P = b.prob[p];
    const i64 c = b.col_first[blockIdx.x] + threadIdx.x;
    const i64 ncol = (i64)st[S_NADD] * P.nd;
    if (c >= ncol) return;
    const int c0 = st[S_C0];
    const i64 v = c / P.nd, J = c - v * P.nd;
    const i64 Ja = J / P.NB, Jb = J - Ja * P.NB;
    const T *X = xcols<T>(b, P, c0) + v * P.nd;
    T *W = wspace<T>(b, P);
    const int2 *pa = b.tab + P.tab_a + Ja * P.nea, *pb = b.tab + P.tab_b + Jb * P.neb;
    int ia = 0, ib = 0;
    int2 ea = pa[0], eb = pb[0];
    for (int q = 0; q < P.npair; ++q) {
        T acc = (T)0;
        if (ia < P.nea && (ea.y & 0xffff) == q) {
            acc += ((ea.y >> 16) ? (T)-1 : (T)1) * X[(i64)ea.x * P.NB + Jb];
            if (++ia < P.nea) ea = pa[ia];
        }
        if (ib < P.neb && (eb.y & 0xffff) == q) {
            acc += ((eb.y >> 16) ? (T)-1 : (T)1) * X[Ja * P.NB + eb.x];
            if (++ib < P.neb) eb = pb[ib];
        }
        W[(i64)q * P.ldw + c] = acc;
    }
}

typedef double bf64x4 __attribute__((ext_vector_type(4)));
typedef float bf32x4 __attribute__((ext_vector_type(4)));

// In-place product E = 1/2 V D + k x on the problem's workspace, 128 columns per
// block (see k_mf_gemm); the row-tile count varies per problem (mt <= 8).
template <typename T>
__global__ void __launch_bounds__(256) k_bmf_gemm(BView b) {
    const int p = b.tile_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!sigma_here<T>(st)) return;
    const BProb P = b.prob[p];
    const i64 ncol = (i64)st[S_NADD] * P.nd;
    const i64 cb = b.tile_first[blockIdx.x];
    if (cb >= ncol) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const i64 c0 = cb + wave * 32;
    const int KP = P.kpad, MT = P.mt;
    const T *V = (sizeof(T) == 8) ? (const T *)(b.V64p + P.vint) : (const T *)(b.V32p + P.vint);
    const T *kv = (sizeof(T) == 8) ? (const T *)(b.k64p + P.kv) : (const T *)(b.k32p + P.kv);
    const T *X = xcols<T>(b, P, b.st[p * ST + S_C0]);
    T *W = wspace<T>(b, P);
    typedef typename std::conditional<sizeof(T) == 8, bf64x4, bf32x4>::type acc_t;
    acc_t acc[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        acc[m][0] = acc_t{};
        acc[m][1] = acc_t{};
    }
    const bool live0 = c0 + lr < ncol, live1 = c0 + 16 + lr < ncol;
    for (int k0 = 0; k0 < KP; k0 += 4) {
        const T b0 = live0 ? W[(i64)(k0 + lk) * P.ldw + c0 + lr] : (T)0;
        const T b1 = live1 ? W[(i64)(k0 + lk) * P.ldw + c0 + 16 + lr] : (T)0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (m < MT) {
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
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const i64 c = c0 + 16 * t + lr;
        const T x = (c < ncol) ? X[c] : (T)0;
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (m < MT)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int q = (sizeof(T) == 8) ? 16 * m + lk + 4 * r : 16 * m + 4 * lk + r;
                    W[(i64)q * P.ldw + c] = (T)0.5 * acc[m][t][r] + kv[q] * x;
                }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bmf_sigma(BView b) {
    const int p = b.col_blk[blockIdx.x];
    const int *st = b.st + p * ST;
    if (!sigma_here<T>(st)) return;
    const BProb P = b.prob[p];
    const i64 c = b.col_first[blockIdx.x] + threadIdx.x;
    const i64 ncol = (i64)st[S_NADD] * P.nd;
    if (c >= ncol) return;
    const int c0 = st[S_C0];
    const i64 v = c / P.nd, I = c - v * P.nd;
    const i64 Ia = I / P.NB, Ib = I - Ia * P.NB;
    const T *E = wspace<T>(b, P) + v * P.nd;
    T acc = (T)0;
    const int2 *pa = b.tab + P.tab_a + Ia * P.nea;
    for (int t = 0; t < P.nea; ++t) {
        const int2 e = pa[t];
        acc += ((e.y >> 16) ? (T)-1 : (T)1) * E[(i64)(e.y & 0xffff) * P.ldw + (i64)e.x * P.NB + Ib];
    }
    const int2 *pb = b.tab + P.tab_b + Ib * P.neb;
    for (int t = 0; t < P.neb; ++t) {
        const int2 e = pb[t];
        acc += ((e.y >> 16) ? (T)-1 : (T)1) * E[(i64)(e.y & 0xffff) * P.ldw + Ia * P.NB + e.x];
    }
    ycols<T>(b, P, c0)[v * P.nd + I] = acc;
}

// Initial block (as the single-problem solver): unit vectors on the lowest
// diagonal entries plus a small fixed perturbation; in a parity sector the
// sector's combination of a determinant and its partner, perturbation projected.
// Start column j of sub-problem p: sa[j] (with weight ca[j]) and sb[j] (cb[j], or
// sb[j] < 0). Row-mapped.
__global__ void __launch_bounds__(256) k_binit(BView b, const i64 *sa, const i64 *sb, const double *ca,
                                               const double *cb) {
    const int p = b.row_blk[blockIdx.x];
    const BProb P = b.prob[p];
    double *V = b.V + P.basis;
    const i64 r0 = b.row_first[blockIdx.x], r1 = min(r0 + (i64)b.RB, P.nd);
    const double amp = 2e-3 * sqrt(3.0 / (double)P.nd);
    const int nstart = b.st[p * ST + S_NADD];
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256)
        for (int j = 0; j < nstart; ++j) {
            const u64 seed = 0x243f6a8885a308d3ULL + (u64)j;
            double x = amp * ((double)(mix64(seed + (u64)r) >> 11) * 0x1p-53 - 0.5);
            if (P.sector != 0) {
                const double xp =
                    amp * ((double)(mix64(seed + (u64)swap_partner(r, P.swap_nb)) >> 11) * 0x1p-53 - 0.5);
                x = 0.5 * (x + P.sector * xp);
            }
            const i64 q = (i64)p * NB2 + j;
            if (r == sa[q]) x += ca[q];
            if (r == sb[q]) x += cb[q];
            V[(i64)j * P.nd + r] = x;
        }
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
struct Batch {
    pyci_ctx *ctx = nullptr;
    hipStream_t s = nullptr;  // the stream the batch's work and graph go to
    int nprob = 0, k = 0, kb0 = 0, mmax = 0, S = 0, RB = 1024;
    std::vector<BProb> hp;
    DevBuf<BProb> prob;
    DevBuf<int2> tab;
    DevBuf<double> V64p, k64p, diag, V, W, X, WX, T, work64, theta, Y, G, res, best, out, C, Rinv, nrm0, energy,
        partial, one, two;
    DevBuf<float> V32p, k32p, S32x, S32y, work32;
    DevBuf<int> st, sel, row_blk, col_blk, tile_blk, nactive;
    DevBuf<i64> row_first, col_first, tile_first;
    DevBuf<unsigned long long> tconv;
    i64 nrowblk = 0, ncolblk = 0, ntile = 0;
    BView view{};
    // large sub-problems: the fused matrix-free pass (direct.hip), one launch pair per
    // sub-problem and precision in the captured iteration, gated by a device descriptor
    std::vector<int> bigp;     // sub-problem indices
    std::vector<MfView> bigv;  // their operator views
    DevBuf<int> bigp_d;
    DevBuf<int2> mtab;         // merged tables with sentinels
    DevBuf<long long> dyn;     // per big sub-problem: {nvec, X, Y} FP64, then FP32
    DevBuf<unsigned char> fws; // shared E workspace of the fused pass
};

// Sub-problems with at least this many determinants use the fused pass.
constexpr i64 kBigND = 65536;

// The device descriptors of the fused-pass launches: the new columns of an active
// sub-problem in the precision of its phase (offsets into V / W, or the FP32 copies).
__global__ void k_btask(BView b, const int *bigp, int nbig, long long *dyn) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nbig) return;
    const int p = bigp[q];
    const int *st = b.st + p * ST;
    const BProb P = b.prob[p];
    const bool on = st[S_ACTIVE] && st[S_NADD] > 0;
    long long *d64 = dyn + 6 * q, *d32 = d64 + 3;
    const long long off = P.basis + (long long)st[S_C0] * P.nd;
    d64[0] = (on && st[S_PHASE] == 0) ? st[S_NADD] : 0;
    d64[1] = off;
    d64[2] = off;
    d32[0] = (on && st[S_PHASE] == 1) ? st[S_NADD] : 0;
    d32[1] = P.s32;
    d32[2] = P.s32;
}

static void build_row_map(Batch &B, int RB) {
    std::vector<int> blk;
    std::vector<i64> first;
    for (int p = 0; p < B.nprob; ++p) {
        B.hp[(size_t)p].rb0 = (i64)blk.size();
        for (i64 r = 0; r < B.hp[(size_t)p].nd; r += RB) {
            blk.push_back(p);
            first.push_back(r);
        }
        B.hp[(size_t)p].nrb = (i64)blk.size() - B.hp[(size_t)p].rb0;
    }
    B.nrowblk = (i64)blk.size();
    B.row_blk.alloc(blk.size());
    B.row_first.alloc(first.size());
    hipStream_t s = B.s;
    PE_HIP(hipMemcpyAsync(B.row_blk.p, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.row_first.p, first.data(), sizeof(i64) * first.size(), hipMemcpyHostToDevice, s));
    B.partial.alloc((size_t)B.nrowblk * B.S);
    PE_HIP(hipMemcpyAsync(B.prob.p, B.hp.data(), sizeof(BProb) * B.hp.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipStreamSynchronize(s));
    B.RB = RB;
    B.view.RB = RB;
    B.view.row_blk = B.row_blk.p;
    B.view.row_first = B.row_first.p;
    B.view.partial = B.partial.p;
}

static void launch_rows(Batch &B, void (*kern)(BView), hipStream_t s) {
    kern<<<(unsigned)B.nrowblk, 256, 0, s>>>(B.view);
    PE_LAUNCH_CHECK();
}

static void orthonormalize_and_apply(Batch &B, hipStream_t s) {
    const BView &v = B.view;
    const unsigned np = (unsigned)B.nprob;
    const unsigned gy_full = (unsigned)(((B.mmax + 7) / 8) * ((B.kb0 + 7) / 8));
    const unsigned gy_new = (unsigned)(((B.kb0 + 7) / 8) * ((B.kb0 + 7) / 8));
    const unsigned ry = (unsigned)((B.S + 3) / 4);
    // norms of the new columns before orthogonalisation
    k_bgram<<<dim3((unsigned)B.nrowblk, gy_new), 256, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    k_bred<<<dim3(np, ry), 256, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    k_bchol<<<(np + 63) / 64, 64, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    for (int pass = 0; pass < 2; ++pass) {  // CGS2 against the basis
        k_bgram<<<dim3((unsigned)B.nrowblk, gy_full), 256, 0, s>>>(v, 0);
        PE_LAUNCH_CHECK();
        k_bred<<<dim3(np, ry), 256, 0, s>>>(v, 0);
        PE_LAUNCH_CHECK();
        launch_rows(B, k_bupd, s);
    }
    // Cholesky-QR of the new block, dropping dependent columns
    k_bgram<<<dim3((unsigned)B.nrowblk, gy_new), 256, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    k_bred<<<dim3(np, ry), 256, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    k_bchol<<<(np + 63) / 64, 64, 0, s>>>(v, 0);
    PE_LAUNCH_CHECK();
    launch_rows(B, k_bapplyR, s);
    // sigma of the new columns: FP64 problems, then FP32 ones; small sub-problems in
    // the batched kernels (one launch each for all), large ones by the fused pass
    const int nbig = (int)B.bigp.size();
    if (nbig) {
        k_btask<<<(unsigned)((nbig + 63) / 64), 64, 0, s>>>(v, B.bigp_d.p, nbig, B.dyn.p);
        PE_LAUNCH_CHECK();
    }
    if (B.ncolblk) {
        k_bmf_d<double><<<(unsigned)B.ncolblk, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
        k_bmf_gemm<double><<<(unsigned)B.ntile, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
        k_bmf_sigma<double><<<(unsigned)B.ncolblk, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
    }
    for (int i = 0; i < nbig; ++i) {
        const BProb &P = B.hp[(size_t)B.bigp[(size_t)i]];
        mf_sigma_batched<double>(B.bigv[(size_t)i], P.nd, B.kb0, B.V.p, B.W.p, (double *)B.fws.p, B.V64p.p + P.vint,
                                 B.k64p.p + P.kv, B.dyn.p + 6 * i, P.sector, s);
    }
    k_bcvt<<<(unsigned)B.nrowblk, 256, 0, s>>>(v, 0);
    PE_LAUNCH_CHECK();
    if (B.ncolblk) {
        k_bmf_d<float><<<(unsigned)B.ncolblk, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
        k_bmf_gemm<float><<<(unsigned)B.ntile, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
        k_bmf_sigma<float><<<(unsigned)B.ncolblk, 256, 0, s>>>(v);
        PE_LAUNCH_CHECK();
    }
    for (int i = 0; i < nbig; ++i) {
        const BProb &P = B.hp[(size_t)B.bigp[(size_t)i]];
        mf_sigma_batched<float>(B.bigv[(size_t)i], P.nd, B.kb0, B.S32x.p, B.S32y.p, (float *)B.fws.p,
                                B.V32p.p + P.vint, B.k32p.p + P.kv, B.dyn.p + 6 * i + 3, P.sector, s);
    }
    k_bcvt<<<(unsigned)B.nrowblk, 256, 0, s>>>(v, 1);
    PE_LAUNCH_CHECK();
    // Rayleigh-matrix entries of the new columns
    k_bgram<<<dim3((unsigned)B.nrowblk, gy_full), 256, 0, s>>>(v, 2);
    PE_LAUNCH_CHECK();
    k_bred<<<dim3(np, ry), 256, 0, s>>>(v, 2);
    PE_LAUNCH_CHECK();
    k_bgupd<<<(np + 63) / 64, 64, 0, s>>>(v);
    PE_LAUNCH_CHECK();
}

static void iteration(Batch &B, hipStream_t s) {
    const BView &v = B.view;
    k_bjac<<<(unsigned)B.nprob, 64, 0, s>>>(v);
    PE_LAUNCH_CHECK();
    launch_rows(B, k_britz, s);
    k_bctl<<<(unsigned)B.nprob, 256, 0, s>>>(v);
    PE_LAUNCH_CHECK();
    launch_rows(B, k_bmove, s);
    orthonormalize_and_apply(B, s);
    k_bcount<<<1, 256, 0, s>>>(v);
    PE_LAUNCH_CHECK();
}

// The replayed-graph path over all given problems, on stream s.
static void batch_graph(pyci_ctx *ctx, hipStream_t s, i64 nprob, const i64 *nbasis, const i64 *nup, const i64 *ndn,
                        const double *ecore, const double *const *one_mo, const double *const *two_mo, i64 nroot,
                        double tol, int precision, i64 maxiter, double *evals, double *latency_ms, i64 *iters,
                        i64 *nsigma, int *converged, pyci_batch_stats *stats) {
    const auto t_start = std::chrono::steady_clock::now();
    Batch B;
    B.ctx = ctx;
    B.s = s;
    B.k = (int)nroot;
    B.mmax = (int)std::max<i64>({2 * nroot + 1, 10, 6 * nroot});
    PE_REQUIRE(B.mmax <= BM_MAX, PYCI_EUNSUPPORTED, "subspace too large");
    B.kb0 = (B.k == 1) ? 1 : std::min({2 * B.k, NB2, B.mmax - B.k});
    B.S = std::max(B.mmax * B.k, B.kb0 * B.kb0);
    std::vector<i64> binom(65 * 65);
    for (int a = 0; a <= 64; ++a)
        for (int c = 0; c <= 64; ++c) binom[(size_t)a * 65 + c] = binom_host(a, c);
    // per caller problem: tables, pair integrals, diagonal slot (shared by its sectors)
    struct Orig {
        int n, na, nb;
        i64 nd, NB, tab_a, tab_b, vint, kv, vec, one, two;
        int npair, kpad, nea, neb;
    };
    std::vector<Orig> orig((size_t)nprob);
    std::vector<int2> tabs;
    std::vector<double> V64, k64;
    i64 vec = 0, one_off = 0, two_off = 0;
    for (i64 p = 0; p < nprob; ++p) {
        Orig &O = orig[(size_t)p];
        O.n = (int)nbasis[p];
        O.na = (int)nup[p];
        O.nb = (int)ndn[p];
        PE_REQUIRE(O.n >= 2 && O.n <= 15 && O.na >= 1 && O.nb >= 1 && O.na <= O.n && O.nb <= O.n && O.na >= O.nb,
                   PYCI_EUNSUPPORTED, "batched problems need 2 <= nbasis <= 15 and 1 <= ndn <= nup <= nbasis");
        const std::vector<u64> sa = colex_strings(O.n, O.na), sb = colex_strings(O.n, O.nb);
        O.NB = (i64)sb.size();
        O.nd = (i64)sa.size() * O.NB;
        PE_REQUIRE(O.nd >= B.mmax, PYCI_EUNSUPPORTED, "a batched problem is smaller than the Davidson subspace");
        O.npair = O.n * (O.n + 1) / 2;
        O.kpad = (O.npair + 15) / 16 * 16;
        PE_REQUIRE(O.kpad / 16 <= 8, PYCI_EUNSUPPORTED, "pair dimension too large");
        O.nea = O.na * (O.n - O.na + 1);
        O.neb = O.nb * (O.n - O.nb + 1);
        O.tab_a = (i64)tabs.size();
        const auto ta = string_table(O.n, sa, O.nea, binom);
        tabs.insert(tabs.end(), ta.begin(), ta.end());
        O.tab_b = (i64)tabs.size();
        const auto tb = string_table(O.n, sb, O.neb, binom);
        tabs.insert(tabs.end(), tb.begin(), tb.end());
        std::vector<double> Vp, kp;
        pair_integrals(O.n, one_mo[p], two_mo[p], O.kpad, Vp, kp);
        O.vint = (i64)V64.size();
        V64.insert(V64.end(), Vp.begin(), Vp.end());
        O.kv = (i64)k64.size();
        k64.insert(k64.end(), kp.begin(), kp.end());
        O.vec = vec;
        vec += O.nd;
        O.one = one_off;
        O.two = two_off;
        one_off += (i64)O.n * O.n;
        two_off += (i64)O.n * O.n * O.n * O.n;
    }
    // sub-problems: both parity sectors of P where it is known and both sectors can
    // hold the subspace comfortably, else the whole space
    i64 basis = 0, ritz = 0, corr = 0, s32 = 0, work = 0;
    std::vector<int> col_blk, tile_blk;
    std::vector<i64> col_first, tile_first;
    for (i64 p = 0; p < nprob; ++p) {
        const Orig &O = orig[(size_t)p];
        const i64 odd_dim = (O.nd - O.NB) / 2;
        const bool split = O.na == O.nb && odd_dim >= 2 * B.mmax && O.nd - odd_dim >= 2 * B.mmax;
        for (int sector : (split ? std::vector<int>{+1, -1} : std::vector<int>{0})) {
            BProb P{};
            P.nd = O.nd;
            P.NB = O.NB;
            P.n = O.n;
            P.npair = O.npair;
            P.kpad = O.kpad;
            P.nea = O.nea;
            P.neb = O.neb;
            P.mt = O.kpad / 16;
            P.tab_a = O.tab_a;
            P.tab_b = O.tab_b;
            P.vint = O.vint;
            P.kv = O.kv;
            P.vec = O.vec;
            P.basis = basis;
            basis += (i64)B.mmax * P.nd;
            P.ritz = ritz;
            ritz += (i64)2 * B.k * P.nd;
            P.corr = corr;
            corr += (i64)B.k * P.nd;
            P.s32 = s32;
            s32 += (i64)B.kb0 * P.nd;
            const i64 cols = (i64)B.kb0 * P.nd;  // widest sigma: the start block
            P.big = P.nd >= kBigND ? 1 : 0;
            const int sp = (int)B.hp.size();
            if (P.big) {
                // the shared workspace of the fused pass instead
                P.ldw = 0;
                P.work = 0;
                B.bigp.push_back(sp);
            } else {
                P.ldw = (cols + 127) / 128 * 128;
                P.work = work;
                work += (i64)P.kpad * P.ldw;
                for (i64 c = 0; c < cols; c += 256) {
                    col_blk.push_back(sp);
                    col_first.push_back(c);
                }
                for (i64 c = 0; c < cols; c += 128) {
                    tile_blk.push_back(sp);
                    tile_first.push_back(c);
                }
            }
            P.ecore = ecore[p];
            P.swap_nb = (sector != 0) ? O.NB : 0;
            P.sector = sector;
            P.orig = (int)p;
            B.hp.push_back(P);
        }
    }
    B.nprob = (int)B.hp.size();
    const int nsub = B.nprob;
    PE_REQUIRE(col_blk.size() < (1u << 31) && tile_blk.size() < (1u << 31), PYCI_EUNSUPPORTED, "batch too large");
    B.prob.alloc((size_t)nsub);
    B.tab.alloc(tabs.size());
    B.V64p.alloc(V64.size());
    B.k64p.alloc(k64.size());
    B.V32p.alloc(V64.size());
    B.k32p.alloc(k64.size());
    std::vector<float> V32(V64.begin(), V64.end()), k32(k64.begin(), k64.end());
    PE_HIP(hipMemcpyAsync(B.tab.p, tabs.data(), sizeof(int2) * tabs.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.V64p.p, V64.data(), sizeof(double) * V64.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.k64p.p, k64.data(), sizeof(double) * k64.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.V32p.p, V32.data(), sizeof(float) * V32.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.k32p.p, k32.data(), sizeof(float) * k32.size(), hipMemcpyHostToDevice, s));
    {
        // fused-pass data of the large sub-problems
        std::vector<int2> mt;
        std::vector<i64> moff;
        size_t ws = 0;
        for (int q : B.bigp) {
            const BProb &P = B.hp[(size_t)q];
            const i64 na = P.nd / P.NB;
            moff.push_back((i64)mt.size());
            mt.insert(mt.end(), tabs.begin() + P.tab_a, tabs.begin() + P.tab_a + na * P.nea);
            for (int t = 0; t < P.neb; ++t)  // beta entries slot-major
                for (i64 r = 0; r < P.NB; ++r) mt.push_back(tabs[(size_t)(P.tab_b + r * P.neb + t)]);
            mt.push_back(int2{0, mf_sink_row(P.kpad)});
            ws = std::max(ws, (size_t)(((P.nd * B.kb0 + 63) / 64) * 64) * (size_t)P.kpad * sizeof(double));
        }
        if (!B.bigp.empty()) {
            B.mtab.alloc(mt.size());
            PE_HIP(hipMemcpyAsync(B.mtab.p, mt.data(), sizeof(int2) * mt.size(), hipMemcpyHostToDevice, s));
            B.fws.alloc(ws);
            B.dyn.alloc(6 * B.bigp.size());
            B.bigp_d.alloc(B.bigp.size());
            PE_HIP(hipMemcpyAsync(B.bigp_d.p, B.bigp.data(), sizeof(int) * B.bigp.size(), hipMemcpyHostToDevice, s));
            for (size_t i = 0; i < B.bigp.size(); ++i) {
                const BProb &P = B.hp[(size_t)B.bigp[i]];
                B.bigv.push_back(MfView{B.mtab.p + moff[i], B.tab.p + P.tab_a, B.tab.p + P.tab_b, P.nd / P.NB, P.NB,
                                        P.npair, P.kpad, P.nea, P.neb});
            }
        }
        // mt leaves scope here: a copy from pageable memory may still be pending
        PE_HIP(hipStreamSynchronize(s));
    }
    B.diag.alloc((size_t)vec);
    B.V.alloc((size_t)basis);
    B.W.alloc((size_t)basis);
    B.X.alloc((size_t)ritz);
    B.WX.alloc((size_t)ritz);
    B.T.alloc((size_t)corr);
    B.S32x.alloc((size_t)s32);
    B.S32y.alloc((size_t)s32);
    B.work64.alloc((size_t)work);
    PE_HIP(hipMemsetAsync(B.work64.p, 0, sizeof(double) * work, s));
    if (precision == PYCI_PREC_MIXED) {
        B.work32.alloc((size_t)work);
        PE_HIP(hipMemsetAsync(B.work32.p, 0, sizeof(float) * work, s));
    } else {
        B.work32.alloc(1);
    }
    // per-sub-problem state
    B.st.alloc((size_t)nsub * ST);
    B.sel.alloc((size_t)nsub * BK_MAX);
    B.theta.alloc((size_t)nsub * BM_MAX);
    B.Y.alloc((size_t)nsub * BM_MAX * 2 * BK_MAX);
    B.G.alloc((size_t)nsub * BM_MAX * BM_MAX);
    B.res.alloc((size_t)nsub * BK_MAX);
    B.best.alloc((size_t)nsub * BK_MAX);
    B.out.alloc((size_t)nsub * B.S);
    B.Rinv.alloc((size_t)nsub * NB2 * NB2);
    B.nrm0.alloc((size_t)nsub * NB2);
    B.energy.alloc((size_t)nsub * BK_MAX);
    B.tconv.alloc((size_t)nsub);
    B.nactive.alloc(1);
    PE_HIP(hipMemsetAsync(B.G.p, 0, sizeof(double) * nsub * BM_MAX * BM_MAX, s));
    PE_HIP(hipMemsetAsync(B.tconv.p, 0, sizeof(unsigned long long) * nsub, s));
    B.col_blk.alloc(col_blk.size());
    B.col_first.alloc(col_first.size());
    B.tile_blk.alloc(tile_blk.size());
    B.tile_first.alloc(tile_first.size());
    PE_HIP(hipMemcpyAsync(B.col_blk.p, col_blk.data(), sizeof(int) * col_blk.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.col_first.p, col_first.data(), sizeof(i64) * col_first.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.tile_blk.p, tile_blk.data(), sizeof(int) * tile_blk.size(), hipMemcpyHostToDevice, s));
    PE_HIP(hipMemcpyAsync(B.tile_first.p, tile_first.data(), sizeof(i64) * tile_first.size(), hipMemcpyHostToDevice,
                          s));
    B.ncolblk = (i64)col_blk.size();
    B.ntile = (i64)tile_blk.size();
    // diagonals (one launch per caller problem at set-up) from device copies of the integrals
    B.one.alloc((size_t)one_off);
    B.two.alloc((size_t)two_off);
    DevBuf<u64> strs;
    for (i64 p = 0; p < nprob; ++p) {
        const Orig &O = orig[(size_t)p];
        PE_HIP(hipMemcpyAsync(B.one.p + O.one, one_mo[p], sizeof(double) * O.n * O.n, hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(B.two.p + O.two, two_mo[p], sizeof(double) * O.n * O.n * O.n * O.n,
                              hipMemcpyHostToDevice, s));
        const std::vector<u64> sa = colex_strings(O.n, O.na), sb = colex_strings(O.n, O.nb);
        strs.alloc(sa.size() + sb.size());
        PE_HIP(hipMemcpyAsync(strs.p, sa.data(), sizeof(u64) * sa.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(strs.p + sa.size(), sb.data(), sizeof(u64) * sb.size(), hipMemcpyHostToDevice, s));
        HamView H{O.n, B.one.p + O.one, B.two.p + O.two};
        mf_diag_launch(O.nd, O.NB, strs.p, strs.p + sa.size(), H, B.diag.p + O.vec, s);
        PE_HIP(hipStreamSynchronize(s));  // strs is reused by the next problem
    }
    // start vectors per sub-problem (host): the lowest diagonal entries, in a sector
    // a determinant with its partner (closed shells only in the even sector)
    DevBuf<i64> dsa, dsb;
    DevBuf<double> dca, dcb;
    std::vector<int> nstart((size_t)nsub, 0);
    {
        std::vector<double> hd((size_t)vec);
        PE_HIP(hipMemcpyAsync(hd.data(), B.diag.p, sizeof(double) * vec, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        std::vector<i64> sa((size_t)nsub * NB2, -1), sb((size_t)nsub * NB2, -1);
        std::vector<double> ca((size_t)nsub * NB2, 0.0), cb((size_t)nsub * NB2, 0.0);
        for (int q = 0; q < nsub; ++q) {
            const BProb &P = B.hp[(size_t)q];
            const double *d = hd.data() + P.vec;
            const i64 want = std::min<i64>(P.nd, 4 * (i64)B.kb0 + 64);
            std::vector<i64> idx((size_t)P.nd);
            std::iota(idx.begin(), idx.end(), 0);
            std::partial_sort(idx.begin(), idx.begin() + want, idx.end(),
                              [&](i64 a, i64 c) { return d[a] < d[c] || (d[a] == d[c] && a < c); });
            std::vector<char> taken;
            if (P.sector != 0) taken.assign((size_t)P.nd, 0);
            int j = 0;
            for (i64 t = 0; t < want && j < B.kb0; ++t) {
                const i64 d0 = idx[(size_t)t];
                const size_t o = (size_t)q * NB2 + j;
                if (P.sector == 0) {
                    sa[o] = d0;
                    ca[o] = 1.0;
                    ++j;
                    continue;
                }
                const i64 d1 = (d0 % P.NB) * P.NB + d0 / P.NB;
                if (taken[(size_t)d0] || (d1 == d0 && P.sector < 0)) continue;
                taken[(size_t)d0] = taken[(size_t)d1] = 1;
                sa[o] = d0;
                if (d1 == d0) {
                    ca[o] = 1.0;
                } else {
                    ca[o] = std::sqrt(0.5);
                    sb[o] = d1;
                    cb[o] = P.sector * std::sqrt(0.5);
                }
                ++j;
            }
            PE_REQUIRE(j >= B.k, PYCI_EUNSUPPORTED, "too few start vectors in a parity sector");
            nstart[(size_t)q] = j;
        }
        dsa.alloc(sa.size());
        dsb.alloc(sb.size());
        dca.alloc(ca.size());
        dcb.alloc(cb.size());
        PE_HIP(hipMemcpyAsync(dsa.p, sa.data(), sizeof(i64) * sa.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(dsb.p, sb.data(), sizeof(i64) * sb.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(dca.p, ca.data(), sizeof(double) * ca.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(dcb.p, cb.data(), sizeof(double) * cb.size(), hipMemcpyHostToDevice, s));
        // sa, sb, ca, cb leave scope here: the copies from them may still be pending
        PE_HIP(hipStreamSynchronize(s));
    }
    {
        std::vector<int> st0((size_t)nsub * ST, 0);
        std::vector<double> best0((size_t)nsub * BK_MAX, 1e300);
        for (int q = 0; q < nsub; ++q) {
            int *w = &st0[(size_t)q * ST];
            w[S_NADD] = nstart[(size_t)q];
            w[S_ACTIVE] = 1;
            w[S_PHASE] = (precision == PYCI_PREC_MIXED) ? 1 : 0;
        }
        PE_HIP(hipMemcpyAsync(B.st.p, st0.data(), sizeof(int) * st0.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(B.best.p, best0.data(), sizeof(double) * best0.size(), hipMemcpyHostToDevice, s));
        PE_HIP(hipStreamSynchronize(s));
    }
    // device view
    BView &v = B.view;
    v.nprob = nsub;
    v.k = B.k;
    v.kb0 = B.kb0;
    v.mmax = B.mmax;
    v.S = B.S;
    v.tol = tol;
    v.switch_tol = 2e-5;
    v.maxiter = (int)((maxiter > 0) ? std::min<i64>(maxiter, 1 << 30) : 3000);
    v.prob = B.prob.p;
    v.col_blk = B.col_blk.p;
    v.col_first = B.col_first.p;
    v.tile_blk = B.tile_blk.p;
    v.tile_first = B.tile_first.p;
    v.tab = B.tab.p;
    v.V64p = B.V64p.p;
    v.k64p = B.k64p.p;
    v.V32p = B.V32p.p;
    v.k32p = B.k32p.p;
    v.diag = B.diag.p;
    v.V = B.V.p;
    v.W = B.W.p;
    v.X = B.X.p;
    v.WX = B.WX.p;
    v.T = B.T.p;
    v.S32x = B.S32x.p;
    v.S32y = B.S32y.p;
    v.work64 = B.work64.p;
    v.work32 = B.work32.p;
    v.st = B.st.p;
    v.sel = B.sel.p;
    v.theta = B.theta.p;
    v.Y = B.Y.p;
    v.G = B.G.p;
    v.res = B.res.p;
    v.best = B.best.p;
    v.out = B.out.p;
    v.Rinv = B.Rinv.p;
    v.nrm0 = B.nrm0.p;
    v.energy = B.energy.p;
    v.tconv = B.tconv.p;
    v.nactive = B.nactive.p;
    // autotune the rows per block of the row-mapped kernels: time the Rayleigh-matrix
    // product and the Ritz pass (their outputs are rewritten before use) per candidate
    int best_rb = 1024;
    {
        build_row_map(B, 1024);
        k_binit<<<(unsigned)B.nrowblk, 256, 0, s>>>(B.view, dsa.p, dsb.p, dca.p, dcb.p);
        PE_LAUNCH_CHECK();
        float best_ms = 1e30f;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        PE_HIP(hipEventCreate(&e0));
        struct EvGuard {
            hipEvent_t &a, &b;
            ~EvGuard() {
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
            }
        } evg{e0, e1};
        PE_HIP(hipEventCreate(&e1));
        for (int rb : {256, 1024, 4096}) {
            build_row_map(B, rb);
            const unsigned gy_full = (unsigned)(((B.mmax + 7) / 8) * ((B.kb0 + 7) / 8));
            k_bgram<<<dim3((unsigned)B.nrowblk, gy_full), 256, 0, s>>>(B.view, 2);  // warm-up
            PE_LAUNCH_CHECK();
            PE_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 3; ++r) {
                k_bgram<<<dim3((unsigned)B.nrowblk, gy_full), 256, 0, s>>>(B.view, 2);
                PE_LAUNCH_CHECK();
                k_britz<<<(unsigned)B.nrowblk, 256, 0, s>>>(B.view);
                PE_LAUNCH_CHECK();
            }
            PE_HIP(hipEventRecord(e1, s));
            PE_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PE_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best_rb = rb;
            }
        }
        build_row_map(B, best_rb);
    }
    // prologue: orthonormalise the initial block and take its sigma
    unsigned long long t0 = 0;
    const auto t_setup = std::chrono::steady_clock::now();
    {
        // a wall-clock stamp on the device for the latencies
        DevBuf<unsigned long long> stamp(1);
        mf_wallclock_launch(stamp.p, s);
        PE_HIP(hipMemcpyAsync(&t0, stamp.p, sizeof(t0), hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
    }
    k_binit<<<(unsigned)B.nrowblk, 256, 0, s>>>(B.view, dsa.p, dsb.p, dca.p, dcb.p);
    PE_LAUNCH_CHECK();
    orthonormalize_and_apply(B, s);
    // the iteration as a captured graph, replayed until no sub-problem is active
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    PE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        iteration(B, s);
    } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(s, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
    }
    PE_HIP(hipStreamEndCapture(s, &graph));
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        throw EngineError(PYCI_EHIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ie));
    }
    i64 launches = 0;
    int nact = nsub;
    try {
        const i64 cap = (i64)v.maxiter + 64;
        while (nact > 0 && launches < cap) {
            for (int r = 0; r < 4; ++r) PE_HIP(hipGraphLaunch(exec, s));
            launches += 4;
            PE_HIP(hipMemcpyAsync(&nact, B.nactive.p, sizeof(int), hipMemcpyDeviceToHost, s));
            PE_HIP(hipStreamSynchronize(s));
        }
    } catch (...) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        throw;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    // results: the k lowest roots over each problem's sub-problems
    std::vector<double> en((size_t)nsub * BK_MAX);
    std::vector<int> stf((size_t)nsub * ST);
    std::vector<unsigned long long> tc((size_t)nsub);
    PE_HIP(hipMemcpyAsync(en.data(), B.energy.p, sizeof(double) * en.size(), hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(stf.data(), B.st.p, sizeof(int) * stf.size(), hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(tc.data(), B.tconv.p, sizeof(unsigned long long) * tc.size(), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    int rate_khz = 0;
    PE_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    const auto t_end = std::chrono::steady_clock::now();
    int nconv = 0;
    for (i64 p = 0; p < nprob; ++p) {
        std::vector<double> roots;
        bool ok = true;
        unsigned long long tmax = 0;
        i64 it = 0, ns = 0;
        for (int q = 0; q < nsub; ++q) {
            if (B.hp[(size_t)q].orig != (int)p) continue;
            const int *w = &stf[(size_t)q * ST];
            ok = ok && !w[S_ACTIVE] && !w[S_FAIL];
            for (i64 j = 0; j < nroot; ++j) roots.push_back(en[(size_t)(q * BK_MAX + j)]);
            tmax = std::max(tmax, tc[(size_t)q]);
            it = std::max<i64>(it, w[S_ITERS]);
            ns += w[S_NSIG];
        }
        std::sort(roots.begin(), roots.end());
        nconv += ok ? 1 : 0;
        // the reference's order: descending for n > 1
        for (i64 j = 0; j < nroot; ++j) evals[p * nroot + j] = roots[(size_t)(nroot - 1 - j)];
        if (latency_ms) latency_ms[p] = (tmax > t0 && rate_khz > 0) ? (double)(tmax - t0) / (double)rate_khz : -1.0;
        if (iters) iters[p] = it;
        if (nsigma) nsigma[p] = ns;
        if (converged) converged[p] = ok ? 1 : 0;
    }
    if (stats) {
        stats->graph_launches = launches;
        stats->rows_per_block = B.RB;
        stats->total_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
        stats->setup_ms = std::chrono::duration<double, std::milli>(t_setup - t_start).count();
        stats->converged = nconv;
    }
}

// One problem by the single-problem solver: the full FullCI space (add_all_dets),
// the matrix-free operator, and the Davidson solve of davidson.hip with its parity
// sectors on their own streams. A run that does not converge reports NaN roots and
// converged = 0, as the graph path reports its unconverged problems.
static void single_solve(pyci_ctx *ctx, int n, int na, int nb, double ecore, const double *one, const double *two,
                         i64 nroot, double tol, int precision, i64 maxiter, double *ev, i64 *it, i64 *ns, int *conv) {
    auto check = [](int rc) {
        if (rc != PYCI_OK) throw EngineError(rc, pyci_last_error());
    };
    const i64 nd = binom_host(n, na) * binom_host(n, nb);
    std::vector<uint64_t> dets((size_t)(2 * nd));
    check(pyci_fill_all_dets(ctx, PYCI_WFN_FULLCI, n, na, nb, dets.data()));
    pyci_ham *h = nullptr;
    pyci_wfn *w = nullptr;
    pyci_op *op = nullptr;
    struct Guard {
        pyci_ham *&h;
        pyci_wfn *&w;
        pyci_op *&op;
        ~Guard() {
            pyci_op_destroy(op);
            pyci_wfn_destroy(w);
            pyci_ham_destroy(h);
        }
    } guard{h, w, op};
    check(pyci_ham_create(ctx, n, ecore, one, two, &h));
    check(pyci_wfn_create(ctx, PYCI_WFN_FULLCI, n, na, nb, nd, dets.data(), &w));
    check(pyci_op_create_ex(h, w, -1, -1, 1, PYCI_STORAGE_AUTO, &op));
    std::vector<double> vecs((size_t)(nroot * nd));
    pyci_solve_stats st{};
    const int rc = pyci_op_solve(op, nroot, nullptr, -1, maxiter, tol, precision, ev, vecs.data(), &st);
    if (rc == PYCI_ENOCONV) {
        for (i64 j = 0; j < nroot; ++j) ev[j] = std::numeric_limits<double>::quiet_NaN();
        *conv = 0;
    } else {
        check(rc);
        *conv = 1;
    }
    *it = st.iterations;
    *ns = st.sigma_calls;
}

// Problems of at least kSingleND determinants take the single-problem solver, on two
// workers that take them from the largest down, on host threads of their own, while the
// smaller ones run together as the replayed graph on a stream of their own. The
// graph's iteration costs the sum of every active problem's work, so a large
// problem in it sets the pace for all; alone, it runs at the single solver's speed
// (FullCI nbasis 12: 0.10-0.18 s for two roots in the single solver, 1.4-2.0 s of
// latency inside the 64-problem graph). Mid-size problems are better off in the
// graph, with its fused pass (8 problems of 108,900 determinants beside 16 of
// 15,876: 358 ms all in the graph, 464 ms with the 8 on the single solver), so the
// threshold lies between that size and 213,444, the next size measured.
constexpr i64 kSingleND = 150000;

void batch_solve(pyci_ctx *ctx, i64 nprob, const i64 *nbasis, const i64 *nup, const i64 *ndn, const double *ecore,
                 const double *const *one_mo, const double *const *two_mo, i64 nroot, double tol, int precision,
                 i64 maxiter, double *evals, double *latency_ms, i64 *iters, i64 *nsigma, int *converged,
                 pyci_batch_stats *stats) {
    PE_REQUIRE(nprob >= 1, PYCI_EINVAL, "empty batch");
    PE_REQUIRE(nroot >= 1 && nroot <= BK_MAX, PYCI_EUNSUPPORTED, "batched solves take 1..8 roots");
    for (i64 p = 0; p < nprob; ++p)
        PE_REQUIRE(nbasis[p] >= 2 && nbasis[p] <= 15 && ndn[p] >= 1 && ndn[p] <= nup[p] && nup[p] <= nbasis[p],
                   PYCI_EUNSUPPORTED, "batched problems need 2 <= nbasis <= 15 and 1 <= ndn <= nup <= nbasis");
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<i64> big, small;
    for (i64 p = 0; p < nprob; ++p) {
        const i64 nd = binom_host((int)nbasis[p], (int)nup[p]) * binom_host((int)nbasis[p], (int)ndn[p]);
        (nd >= kSingleND ? big : small).push_back(p);
    }
    std::stable_sort(big.begin(), big.end(), [&](i64 a, i64 b) {
        return binom_host((int)nbasis[a], (int)nup[a]) * binom_host((int)nbasis[a], (int)ndn[a]) >
               binom_host((int)nbasis[b], (int)nup[b]) * binom_host((int)nbasis[b], (int)ndn[b]);
    });
    if (big.empty()) {
        batch_graph(ctx, ctx->stream, nprob, nbasis, nup, ndn, ecore, one_mo, two_mo, nroot, tol, precision, maxiter,
                    evals, latency_ms, iters, nsigma, converged, stats);
        if (stats) stats->single_path = 0;
        return;
    }
    // the small problems' arguments and outputs, compacted
    const i64 ns_ = (i64)small.size();
    std::vector<i64> snb(ns_), snu(ns_), snd(ns_), sit(ns_), sns(ns_);
    std::vector<double> sec(ns_), sev((size_t)(ns_ * nroot)), slat(ns_);
    std::vector<const double *> sone(ns_), stwo(ns_);
    std::vector<int> sconv(ns_);
    for (i64 i = 0; i < ns_; ++i) {
        const i64 p = small[(size_t)i];
        snb[i] = nbasis[p], snu[i] = nup[p], snd[i] = ndn[p], sec[i] = ecore[p];
        sone[i] = one_mo[p], stwo[i] = two_mo[p];
    }
    std::vector<double> bev((size_t)(big.size() * nroot)), blat(big.size());
    std::vector<i64> bit(big.size()), bns(big.size());
    std::vector<int> bconv(big.size());
    const int device = ctx->device;
    std::exception_ptr err_big;
    // The large problems go to two workers, each pulling the next largest: the second
    // on a context of its own (streams, workspace, solver memory), so that one
    // solve's kernels fill the GPU while the other waits for its host's Rayleigh-Ritz
    // steps (the problems of ~200,000 determinants leave it idle otherwise).
    std::atomic<size_t> next{0};
    unsigned long long t0 = 0;
    int rate_khz = 0;
    auto worker = [&](pyci_ctx *c) {
        DevBuf<unsigned long long> stamp(1);
        auto now = [&] {
            unsigned long long t = 0;
            mf_wallclock_launch(stamp.p, c->stream);
            PE_HIP(hipMemcpyAsync(&t, stamp.p, sizeof(t), hipMemcpyDeviceToHost, c->stream));
            PE_HIP(hipStreamSynchronize(c->stream));
            return t;
        };
        for (size_t i = next++; i < big.size(); i = next++) {
            const i64 p = big[i];
            single_solve(c, (int)nbasis[p], (int)nup[p], (int)ndn[p], ecore[p], one_mo[p], two_mo[p], nroot, tol,
                         precision, maxiter, &bev[i * (size_t)nroot], &bit[i], &bns[i], &bconv[i]);
            blat[i] = rate_khz > 0 ? (double)(now() - t0) / (double)rate_khz : -1.0;
        }
    };
    std::thread tb([&] {
        pyci_ctx *c2 = nullptr;
        std::exception_ptr err2;
        try {
            PE_HIP(hipSetDevice(device));
            PE_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device));
            {
                DevBuf<unsigned long long> stamp(1);
                mf_wallclock_launch(stamp.p, ctx->stream);
                PE_HIP(hipMemcpyAsync(&t0, stamp.p, sizeof(t0), hipMemcpyDeviceToHost, ctx->stream));
                PE_HIP(hipStreamSynchronize(ctx->stream));
            }
            std::thread t2;
            if (big.size() > 1) {
                if (pyci_ctx_create(device, &c2) != PYCI_OK) throw EngineError(PYCI_EHIP, pyci_last_error());
                t2 = std::thread([&] {
                    try {
                        PE_HIP(hipSetDevice(device));
                        worker(c2);
                    } catch (...) {
                        err2 = std::current_exception();
                        next = big.size();  // the other worker stops after its current problem
                    }
                });
            }
            try {
                worker(ctx);
            } catch (...) {
                next = big.size();
                if (t2.joinable()) t2.join();
                throw;
            }
            if (t2.joinable()) t2.join();
        } catch (...) {
            err_big = std::current_exception();
        }
        if (c2) pyci_ctx_destroy(c2);
        if (!err_big && err2) err_big = err2;
    });
    hipStream_t gs = nullptr;
    pyci_batch_stats gst{};
    try {
        if (ns_ > 0) {
            PE_HIP(hipStreamCreateWithFlags(&gs, hipStreamNonBlocking));
            batch_graph(ctx, gs, ns_, snb.data(), snu.data(), snd.data(), sec.data(), sone.data(), stwo.data(), nroot,
                        tol, precision, maxiter, sev.data(), slat.data(), sit.data(), sns.data(), sconv.data(), &gst);
        }
    } catch (...) {
        tb.join();
        if (gs) (void)hipStreamDestroy(gs);
        throw;
    }
    tb.join();
    if (gs) PE_HIP(hipStreamDestroy(gs));
    if (err_big) std::rethrow_exception(err_big);
    i64 nconv = 0;
    auto put = [&](i64 p, const double *ev, double lat, i64 it, i64 nsg, int cv) {
        for (i64 j = 0; j < nroot; ++j) evals[p * nroot + j] = ev[j];
        if (latency_ms) latency_ms[p] = lat;
        if (iters) iters[p] = it;
        if (nsigma) nsigma[p] = nsg;
        if (converged) converged[p] = cv;
        nconv += cv ? 1 : 0;
    };
    for (i64 i = 0; i < ns_; ++i)
        put(small[(size_t)i], &sev[(size_t)(i * nroot)], slat[(size_t)i], sit[(size_t)i], sns[(size_t)i],
            sconv[(size_t)i]);
    for (size_t i = 0; i < big.size(); ++i)
        put(big[i], &bev[i * (size_t)nroot], blat[i], bit[i], bns[i], bconv[i]);
    if (stats) {
        *stats = gst;
        stats->converged = nconv;
        stats->single_path = (i64)big.size();
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
}

}  // namespace pe
