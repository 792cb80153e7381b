// Block Davidson eigensolver on the GPU.
//
// Reference: SparseOp::solve_ci (sparseop.cpp:114-146) runs Spectra's implicitly
// restarted Lanczos (SymEigsSolver, SmallestAlge) on the host. This solver keeps
// the basis V, its image W = H V, the Ritz vectors and the corrections in HBM
// (column-major, one vector contiguous). Per iteration:
//   sigma     W[:, new] = H V[:, new]            (SpMM, HBM-bound; FP64 or FP32 values)
//   project   G[:, new] = V^T W[:, new]          (GPU reduction, FP64)
//   RR        (theta, Y) = eig(G)                (m x m, m <= ncv, host Jacobi, FP64)
//   residual  X = V Y, R = W Y - X theta, |R|    (one fused GPU pass, FP64)
//   correct   T = R / (theta - diag(H))          (Davidson diagonal preconditioner)
//   orthog.   T -= V (V^T T), twice; normalise   (classical Gram-Schmidt with re-orthogonalisation)
// A root is converged when |r| <= tol * max(|theta|, eps^(2/3)), the criterion
// Spectra applies to its Ritz estimates. In the mixed mode the first phase uses
// the FP32-value sigma with FP32 accumulation; it hands its Ritz vectors to an
// FP64 phase (the residual-correction pass) that converges to the FP64 tolerance.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <numeric>
#include <thread>

#include "internal.hpp"

namespace pe {

constexpr int DT = 256;

// Sums over the per-block partials of an earlier kernel (partial[b * cnt + t],
// b < npart), computed by every block of a consuming kernel into shared memory
// out[t] (t < cnt): a fixed order, so all blocks (and launches) agree bit for bit.
__device__ void reduce_partials_block(const double *partial, int npart, int cnt, double *out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int t = wave; t < cnt; t += nw) {
        double v = 0.0;
        for (int b = lane; b < npart; b += 64) v += partial[(i64)b * cnt + t];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) out[t] = v;
    }
    __syncthreads();
}

// Sum of v over the block, written by thread 0 to *dst.
__device__ inline void block_sum_to(double v, double *dst) {
    __shared__ double red[DT / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < DT / 64; ++w) t += red[w];
        *dst = t;
    }
}

// out[q] = sum over b < npart of p[q][b] for N arrays of one value per block, all
// threads of the block taking part (a fixed order: every block gets the same bits).
template <int N>
__device__ void sum_partials(const double *const (&p)[N], int npart, double *out) {
    __shared__ double red[DT / 64][N];
    double v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < npart; b += DT)
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] += p[q][b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
        if (lane == 0) red[wave][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0.0;
        for (int w = 0; w < DT / 64; ++w) t += red[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

// Second classical Gram-Schmidt pass only where the first removed more than half
// of the squared norm (|t1|^2 < |t0|^2 / 2): "twice is enough" with the
// Daniel-Gragg-Kaufman-Stewart test, decided on the device from the partial norms
// of the copy (p0) and of the first pass (p1).
__device__ bool second_pass(const double *p0, const double *p1, int npart) {
    __shared__ double nn[2];
    const double *const pp[2] = {p0, p1};
    sum_partials<2>(pp, npart, nn);
    return nn[1] < 0.5 * nn[0];
}

// Sums of acc[i] (i < M) over the block to out[i].
template <int M>
__device__ void block_sums_to(double (&acc)[M], double *out) {
    __shared__ double red[DT / 64][M];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double v = acc[i];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < M) {
        double v = 0.0;
        for (int w = 0; w < DT / 64; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// partial[b][ma][mb] = sum over rows of block b of A[:, i] * B[:, j] (8x8 tile per
// blockIdx.y). With gate0 set the launch is the second Gram-Schmidt pass and does
// nothing when the first pass sufficed.
__global__ void __launch_bounds__(DT) k_gram(i64 nd, const double *A, i64 lda, int ma, const double *B, i64 ldb,
                                             int mb, double *partial, const double *gate0, const double *gate1) {
    if (gate0 && !second_pass(gate0, gate1, gridDim.x)) return;
    const int tiles_j = (mb + 7) / 8;
    const int i0 = (blockIdx.y / tiles_j) * 8, j0 = (blockIdx.y % tiles_j) * 8;
    const int ni = min(8, ma - i0), nj = min(8, mb - j0);
    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
    for (i64 r = (i64)blockIdx.x * DT + threadIdx.x; r < nd; r += (i64)gridDim.x * DT) {
        double a[8], b[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = (i < ni) ? A[(i64)(i0 + i) * lda + r] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = (j < nj) ? B[(i64)(j0 + j) * ldb + r] : 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] += a[i] * b[j];
    }
    __shared__ double red[DT / 64][64];
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
        double v = 0.0;
        for (int w = 0; w < DT / 64; ++w) v += red[w][threadIdx.x];
        if (i < ni && j < nj) partial[((i64)blockIdx.x * ma + i0 + i) * mb + j0 + j] = v;
    }
}

// out[t] = sum over b of partial[b][t]; one 256-thread block per output element.
__global__ void __launch_bounds__(256) k_reduce_partials(const double *partial, int nblk, int count, double *out) {
    const int t = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) v += partial[(i64)b * count + t];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[t] = red[0] + red[1] + red[2] + red[3];
}

// T[:, j] += alpha * V[:, 0:m] C[:, j]   (C is m x k column-major on the device)
__global__ void k_gemm_update(i64 nd, const double *V, i64 ldv, int m, const double *C, int k, double alpha,
                              double *T, i64 ldt) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x) {
        for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc += V[(i64)i * ldv + r] * C[j * m + i];
            T[(i64)j * ldt + r] += alpha * acc;
        }
    }
}

// T[:, j] -= V[:, 0:m] C[:, j] with C row-major (C[i * k + j]), as the Gram products leave it
__global__ void k_gemm_sub_rm(i64 nd, const double *V, int m, const double *C, int k, double *T) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x) {
        for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc += V[(i64)i * nd + r] * C[i * k + j];
            T[(i64)j * nd + r] -= acc;
        }
    }
}

// T[:, j] = V[:, 0:m] C[:, j]
__global__ void k_gemm_set(i64 nd, const double *V, i64 ldv, int m, const double *C, int k, double *T, i64 ldt) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x) {
        for (int j = 0; j < k; ++j) {
            double acc = 0.0;
            for (int i = 0; i < m; ++i) acc += V[(i64)i * ldv + r] * C[j * m + i];
            T[(i64)j * ldt + r] = acc;
        }
    }
}

// t -= V[:, 0:m] c with c the sums of cpart (partials of k_gram, m per block),
// and the partial squared norms of the result per block; a gated launch (second
// Gram-Schmidt pass) does nothing when the first pass sufficed. m <= 128.
__global__ void __launch_bounds__(DT) k_update_norm(i64 nd, const double *V, int m, const double *cpart, double *t,
                                                    double *npart, const double *gate0, const double *gate1) {
    if (gate0 && !second_pass(gate0, gate1, gridDim.x)) return;
    __shared__ double c[128];
    reduce_partials_block(cpart, gridDim.x, m, c);
    double ss = 0.0;
    for (i64 r = (i64)blockIdx.x * DT + threadIdx.x; r < nd; r += (i64)gridDim.x * DT) {
        double acc = 0.0;
        for (int i = 0; i < m; ++i) acc += V[(i64)i * nd + r] * c[i];
        const double x = t[r] - acc;
        t[r] = x;
        ss += x * x;
    }
    block_sum_to(ss, npart + blockIdx.x);
}

// Basis-size-specialised forms of the one-column passes (M = the number of basis
// columns, 1..16): every thread handles two rows 256 apart per step and issues all
// their loads before the arithmetic. Columns are nd apart.

// partial[b * M + i] = sum over the rows of block b of A[:, i] x; gated like
// k_gram. With hnb > 0 (nd = hnb^2, the columns of A and x in one parity sector of
// the alpha <-> beta exchange P: z(P r) = +-z(r) with the same sign for all) only the
// rows (Ia, Ib) = (r / hnb, r % hnb) with Ia <= Ib are read; a row with Ia < Ib also
// stands for its partner (Ib, Ia) and counts twice.
template <int M>
__global__ void __launch_bounds__(DT) k_dots(i64 nd, const double *__restrict__ A, const double *__restrict__ x,
                                             double *partial, const double *gate0, const double *gate1,
                                             unsigned hnb) {
    if (gate0 && !second_pass(gate0, gate1, gridDim.x)) return;
    double acc[M];
#pragma unroll
    for (int i = 0; i < M; ++i) acc[i] = 0.0;
    const i64 step = (i64)gridDim.x * 2 * DT;
    for (i64 r0 = (i64)blockIdx.x * 2 * DT + threadIdx.x; r0 < nd; r0 += step) {
        double wt[2], xr[2], a[2][M];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const i64 r = r0 + u * DT;
            wt[u] = (r < nd) ? 1.0 : 0.0;
            if (hnb && r < nd) {
                const unsigned ia = (unsigned)r / hnb, ib = (unsigned)r - ia * hnb;
                wt[u] = ia < ib ? 2.0 : ia == ib ? 1.0 : 0.0;
            }
            xr[u] = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) a[u][i] = 0.0;
            if (wt[u] != 0.0) {
                xr[u] = x[r];
#pragma unroll
                for (int i = 0; i < M; ++i) a[u][i] = A[(i64)i * nd + r];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int i = 0; i < M; ++i) acc[i] += a[u][i] * (wt[u] * xr[u]);
    }
    block_sums_to<M>(acc, partial + (i64)blockIdx.x * M);
}

// t -= V[:, 0:M] c, c the sums of cpart (M per block), and the partial squared
// norms of the result per block; gated like k_update_norm. With rout, block 0 also
// sums the one-value-per-block partials rpart into *rout (the residual norm of the
// Ritz pass, so that it needs no launch of its own).
template <int M>
__global__ void __launch_bounds__(DT) k_update_norm_m(i64 nd, const double *__restrict__ V,
                                                      const double *cpart, double *t, double *npart,
                                                      const double *gate0, const double *gate1,
                                                      const double *rpart, double *rout) {
    if (gate0 && !second_pass(gate0, gate1, gridDim.x)) return;
    if (rout && blockIdx.x == 0) {
        __shared__ double rs[1];
        const double *const rp[1] = {rpart};
        sum_partials<1>(rp, gridDim.x, rs);
        if (threadIdx.x == 0) *rout = rs[0];
    }
    __shared__ double c[M];
    reduce_partials_block(cpart, gridDim.x, M, c);
    double cr[M];
#pragma unroll
    for (int i = 0; i < M; ++i) cr[i] = c[i];
    double ss = 0.0;
    const i64 step = (i64)gridDim.x * 2 * DT;
    for (i64 r0 = (i64)blockIdx.x * 2 * DT + threadIdx.x; r0 < nd; r0 += step) {
        double a[2][M], tv[2];
        i64 r[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            r[u] = min(r0 + u * DT, nd - 1);
            tv[u] = t[r[u]];
#pragma unroll
            for (int i = 0; i < M; ++i) a[u][i] = V[(i64)i * nd + r[u]];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) acc += a[u][i] * cr[i];
            const double x = tv[u] - acc;
            if (r0 + u * DT < nd) {
                t[r[u]] = x;
                ss += x * x;
            }
        }
    }
    block_sum_to(ss, npart + blockIdx.x);
}

// One root, the speculative iteration: with the Ritz coefficients y (M) and theta
// (kernel arguments),
// per row x = V y, wx = W y, res = wx - theta x and the correction
// t = res / (theta - d) (d the diagonal, the denominator kept off zero as in
// k_ritz), written to T. Per block: the squared residual norm (pres), the squared
// norm of t (pnorm) and the first Gram-Schmidt coefficients V^T t (pdots, M per
// block). With the basis, its images and d all in one parity sector (d symmetric
// under P), t is in that sector to the last bit, with no projection.
struct RitzCoef {
    double y[16];  // Ritz coefficients of the root (the first M used)
    double theta;
};

template <int M>
__global__ void __launch_bounds__(DT) k_ritz1(i64 nd, const double *__restrict__ V, const double *__restrict__ W,
                                              RitzCoef cf, const double *__restrict__ diag, double *__restrict__ T,
                                              double *pres, double *pnorm, double *pdots) {
    double y[M];
#pragma unroll
    for (int i = 0; i < M; ++i) y[i] = cf.y[i];
    const double th = cf.theta;
    double acc[M + 2];
#pragma unroll
    for (int i = 0; i < M + 2; ++i) acc[i] = 0.0;
    const i64 step = (i64)gridDim.x * 2 * DT;
    for (i64 r0 = (i64)blockIdx.x * 2 * DT + threadIdx.x; r0 < nd; r0 += step) {
        double v[2][M], w[2][M], d[2];
        i64 r[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            r[u] = min(r0 + u * DT, nd - 1);
            d[u] = diag[r[u]];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                v[u][i] = V[(i64)i * nd + r[u]];
                w[u][i] = W[(i64)i * nd + r[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (r0 + u * DT >= nd) break;
            double x = 0.0, wx = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                x += v[u][i] * y[i];
                wx += w[u][i] * y[i];
            }
            const double res = wx - th * x;
            double den = th - d[u];
            if (fabs(den) < 1e-8) den = (den < 0.0) ? -1e-8 : 1e-8;
            const double t = res / den;
            T[r[u]] = t;
            acc[M] += res * res;
            acc[M + 1] += t * t;
#pragma unroll
            for (int i = 0; i < M; ++i) acc[i] += v[u][i] * t;
        }
    }
    __shared__ double out[M + 2];
    block_sums_to<M + 2>(acc, out);
    __syncthreads();
    if (threadIdx.x < M) pdots[(i64)blockIdx.x * M + threadIdx.x] = out[threadIdx.x];
    if (threadIdx.x == 0) {
        pres[blockIdx.x] = out[M];
        pnorm[blockIdx.x] = out[M + 1];
    }
}

// d[r] = (diag[r] + diag[P r]) / 2 for P's index map (r = Ia nb + Ib -> Ib nb + Ia):
// the same bits at r and P r.
__global__ void k_symdiag(const double *diag, double *d, i64 nd, i64 nb) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x) {
        const i64 a = r / nb, b = r - a * nb;
        d[r] = 0.5 * (diag[r] + diag[b * nb + a]);
    }
}

// dst = src, or dst = (src + sgn P src) / 2 with P the alpha <-> beta exchange of
// a full FullCI space in colex order with nup == ndn (index Ja * nb + Jb <->
// Jb * nb + Ja) when sgn != 0; the partial squared norms of dst per block.
__global__ void __launch_bounds__(DT) k_put(const double *src, double *dst, i64 nd, i64 nb, double sgn,
                                            double *npart) {
    double ss = 0.0;
    for (i64 J = (i64)blockIdx.x * DT + threadIdx.x; J < nd; J += (i64)gridDim.x * DT) {
        double x = src[J];
        if (sgn != 0.0) {
            const i64 a = J / nb, b = J - a * nb;
            x = 0.5 * (x + sgn * src[b * nb + a]);
        }
        dst[J] = x;
        ss += x * x;
    }
    block_sum_to(ss, npart + blockIdx.x);
}

// Ritz vectors and images (only when write_xw), preconditioned residuals and the
// squared residual norms (completed in-kernel into res).
__global__ void __launch_bounds__(DT) k_ritz(i64 nd, const double *V, const double *W, i64 ld, int m,
                                             const double *Y, const double *theta, int k, const double *diag,
                                             int write_xw, double *X, double *WX, double *T, double *npart) {
    double ss[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) ss[j] = 0.0;
    for (i64 r = (i64)blockIdx.x * DT + threadIdx.x; r < nd; r += (i64)gridDim.x * DT) {
        double x[16], w[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = w[j] = 0.0;
        for (int i = 0; i < m; ++i) {
            const double v = V[(i64)i * ld + r], ww = W[(i64)i * ld + r];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < k) {
                    const double y = Y[j * m + i];
                    x[j] += v * y;
                    w[j] += ww * y;
                }
        }
        const double d = diag[r];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < k) {
                const double res = w[j] - theta[j] * x[j];
                double den = theta[j] - d;
                if (fabs(den) < 1e-8) den = (den < 0.0) ? -1e-8 : 1e-8;
                if (write_xw) {
                    X[(i64)j * ld + r] = x[j];
                    WX[(i64)j * ld + r] = w[j];
                }
                T[(i64)j * ld + r] = res / den;
                ss[j] += res * res;
            }
    }
    __shared__ double red[DT / 64][16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        double v = ss[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)k) {
        double v = 0.0;
        for (int w = 0; w < DT / 64; ++w) v += red[w][threadIdx.x];
        npart[(i64)blockIdx.x * k + threadIdx.x] = v;
    }
}

__global__ void k_scale(double *x, i64 n, double s) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) x[i] *= s;
}

// Normalises x by the norm whose square is nn[1]; a vector whose norm fell below
// 1e-10 of its norm before orthogonalisation (square in nn[0]) is numerically
// dependent on the basis and is zeroed instead (the host drops it later).
__device__ inline double checked_scale(double n0sq, double n1sq) {
    const double n0 = sqrt(fmax(n0sq, 0.0)), n1 = sqrt(fmax(n1sq, 0.0));
    return (n1 > 1e-10 * fmax(n0, 1e-300) && n1 > 1e-300) ? 1.0 / n1 : 0.0;
}

__global__ void k_scale_checked(double *x, i64 n, const double *nn) {
    const double s = checked_scale(nn[0], nn[1]);
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) x[i] *= s;
}

// Final normalisation of an orthogonalised correction from the partial squared
// norms (npart per block) before Gram-Schmidt (p0), after the first pass (p1) and
// after the second (p2, used when `gated` and the second pass ran). Writes the
// pair (before, final) to pair for the host.
__global__ void __launch_bounds__(DT) k_scale_final(double *x, i64 n, const double *p0, const double *p1,
                                                    const double *p2, int npart, int gated, double *pair) {
    __shared__ double nn[3];
    const double *const pp[3] = {p0, p1, gated ? p2 : p1};
    sum_partials<3>(pp, npart, nn);
    const double nf = (gated && nn[1] < 0.5 * nn[0]) ? nn[2] : nn[1];
    const double s = checked_scale(nn[0], nf);
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) x[i] *= s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pair[0] = nn[0];
        pair[1] = nf;
    }
}

// dst = (src + s P src) / 2 (see k_put), without the norm.
__global__ void k_parity(const double *src, double *dst, i64 nd, i64 nb, double sgn) {
    for (i64 J = (i64)blockIdx.x * blockDim.x + threadIdx.x; J < nd; J += (i64)gridDim.x * blockDim.x) {
        const i64 a = J / nb, b = J - a * nb;
        dst[J] = 0.5 * (src[J] + sgn * src[b * nb + a]);
    }
}

// Start vector on the device: a perturbation amp (u - 1/2), u uniform in [0, 1)
// from mix64(seed + r) (projected on the parity sector when sgn != 0, partner index
// (r % nb) nb + r / nb), plus w0 at d0 and w1 at d1 (d < 0: none).
__global__ void k_start(double *x, i64 nd, u64 seed, double amp, i64 nb, double sgn, i64 d0, double w0, i64 d1,
                        double w1) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x) {
        // both products rounded on their own (no contraction into the sum), so that
        // v(P r) = sgn v(r) to the last bit
        double v = __dmul_rn(amp, (double)(mix64(seed + (u64)r) >> 11) * 0x1p-53 - 0.5);
        if (sgn != 0.0) {
            const i64 pr = (r % nb) * nb + r / nb;
            v = 0.5 * (v + sgn * __dmul_rn(amp, (double)(mix64(seed + (u64)pr) >> 11) * 0x1p-53 - 0.5));
        }
        if (r == d0) v += w0;
        if (r == d1) v += w1;
        x[r] = v;
    }
}

// col-major (nd x k, ld) <-> row-major (nd x k), with optional precision change
template <typename S, typename D>
__global__ void k_col2row(i64 nd, int k, const S *src, i64 ld, D *dst) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x)
        for (int j = 0; j < k; ++j) dst[r * k + j] = (D)src[(i64)j * ld + r];
}

template <typename S, typename D>
__global__ void k_row2col(i64 nd, int k, const S *src, D *dst, i64 ld) {
    for (i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x; r < nd; r += (i64)gridDim.x * blockDim.x)
        for (int j = 0; j < k; ++j) dst[(i64)j * ld + r] = (D)src[r * k + j];
}

// ---------------------------------------------------------------------------
// Host: cyclic Jacobi eigensolver for the small symmetric Rayleigh matrix.
// ---------------------------------------------------------------------------
static void jacobi_eig(int m, std::vector<double> a, std::vector<double> &w, std::vector<double> &v) {
    // a: m x m row-major symmetric; returns ascending w and eigenvectors as columns of v (row-major m x m)
    v.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) v[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < m; ++p)
            for (int q = 0; q < m; ++q) {
                double x = a[(size_t)p * m + q] * a[(size_t)p * m + q];
                tot += x;
                if (p != q) off += x;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                double apq = a[(size_t)p * m + q];
                if (apq == 0.0) continue;
                double app = a[(size_t)p * m + p], aqq = a[(size_t)q * m + q];
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int r = 0; r < m; ++r) {
                    double arp = a[(size_t)r * m + p], arq = a[(size_t)r * m + q];
                    a[(size_t)r * m + p] = c * arp - s * arq;
                    a[(size_t)r * m + q] = s * arp + c * arq;
                }
                for (int r = 0; r < m; ++r) {
                    double apr = a[(size_t)p * m + r], aqr = a[(size_t)q * m + r];
                    a[(size_t)p * m + r] = c * apr - s * aqr;
                    a[(size_t)q * m + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < m; ++r) {
                    double vrp = v[(size_t)r * m + p], vrq = v[(size_t)r * m + q];
                    v[(size_t)r * m + p] = c * vrp - s * vrq;
                    v[(size_t)r * m + q] = s * vrp + c * vrq;
                }
            }
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return a[(size_t)x * m + x] < a[(size_t)y * m + y]; });
    std::vector<double> vv((size_t)m * m);
    w.resize(m);
    for (int c = 0; c < m; ++c) {
        w[c] = a[(size_t)order[c] * m + order[c]];
        for (int r = 0; r < m; ++r) vv[(size_t)r * m + c] = v[(size_t)r * m + order[c]];
    }
    v.swap(vv);
}

// What the first parity sector's run tells the second one, which runs
// concurrently on another stream and host thread: the first run's state (0
// running, 1 converged, 2 ended unconverged, 3 aborted by an error), its
// completed iterations while it runs, and once it has converged, its highest
// requested root (theta[k - 1]) and iteration count.
struct SectorLink {
    std::atomic<int> state{0};
    std::atomic<i64> done{0};
    double theta = 0.0;
    i64 iters = 0;
};

struct Davidson {
    pyci_op *op;
    hipStream_t s;
    DevBuf<unsigned char> *ws;  // the matrix-free sigma's E workspace (nullptr: the context's)
    i64 nd;
    int k, mmax;
    // working memory: a set kept by the context (or owned, when none is given)
    std::unique_ptr<DavBufs> owned;
    DavBufs &bf;
    DevBuf<double> &V, &W, &X, &WX, &T, &rowbuf, &rowout, &partial, &small;
    DevBuf<double> &RV, &RW;  // thick-restart scratch (nd x 2k each)
    DevBuf<double> &dv;       // small device operands and results (layout: the o* offsets)
    HostBuf<double> &hb;      // page-locked mirror of dv for the per-iteration transfers
    double *hd = nullptr;     // hb as the device sees it: the residual norms, norm pairs and
                              // Rayleigh columns are written there by the kernels themselves
    bool y_dev = false;       // the Ritz coefficients of the last Ritz pass are in dv
    DevBuf<double> &pb;       // per-block partials of the residual norms, the Gram-Schmidt
                              // coefficients and the correction norms (layout: the p* offsets)
    DevBuf<float> &row32, &out32;
    std::vector<double> G;  // mmax x mmax row-major
    int m = 0;
    int nblk;
    int sector = 0;   // 0, or +1 / -1: the basis is kept in that parity sector of P
    i64 swap_nb = 0;  // beta-string count of P's index map
    bool exact = true;          // sigma outputs in the basis' sector to the last bit
    DevBuf<double> &dsym;       // in a sector: the diagonal made symmetric under P
    const double *dg = nullptr; // the diagonal used by the preconditioner
    pyci_solve_stats st{};
    // dv / hb layout (doubles): Ritz coefficients Y (m x k) and theta go down; the
    // squared residual norms, the new G block and the norm pairs of the added
    // columns come back in one copy
    size_t oY, oTH, oYK, oRES, oC1, oC2, oNRM, oPAIR, oG, osz;
    size_t pRES, pN0, pN1, pN2, pC1, pC2;  // in pb: nblk entries per block-partial quantity
    static constexpr int kCgsMax = 128;    // basis size up to which Gram-Schmidt runs on partials

    Davidson(pyci_op *o, int k_, int mmax_, hipStream_t stream = nullptr, DevBuf<unsigned char> *ws_ = nullptr,
             DavBufs *kept = nullptr)
        : op(o), s(stream ? stream : o->ctx->stream), ws(ws_), nd(o->nrow), k(k_), mmax(mmax_),
          owned(kept ? nullptr : new DavBufs), bf(kept ? *kept : *owned), V(bf.V), W(bf.W), X(bf.X), WX(bf.WX),
          T(bf.T), rowbuf(bf.rowbuf), rowout(bf.rowout), partial(bf.partial), small(bf.small), RV(bf.RV),
          RW(bf.RW), dv(bf.dv), hb(bf.hb), pb(bf.pb), row32(bf.row32), out32(bf.out32), dsym(bf.dsym) {
        // a kept set is only written by kernels of this run's stream; work of the
        // previous run that used it was on the same stream
        V.ensure((size_t)(nd * mmax));
        W.ensure((size_t)(nd * mmax));
        X.ensure((size_t)(nd * k));
        WX.ensure((size_t)(nd * k));
        T.ensure((size_t)(nd * k));
        // staging for sigma of up to max(k, start block) vectors at once
        const int kcols = std::max(k, std::min(2 * k, 16));
        rowbuf.ensure((size_t)(nd * kcols));
        rowout.ensure((size_t)(nd * kcols));
        nblk = (int)std::min<i64>(512, (nd + DT - 1) / DT);
        partial.ensure((size_t)nblk * (size_t)std::max(mmax * std::max(k, kcols), 64) + 64);
        small.ensure((size_t)(mmax * mmax + 4 * mmax + 64));
        const size_t M = (size_t)mmax;
        oY = 0;
        oTH = oY + M * 16;
        oYK = oTH + 16;
        oRES = oYK + M * 32;
        oC1 = oRES + 16;
        oC2 = oC1 + M;
        oNRM = oC2 + M;
        oPAIR = oNRM + 3 * 16;
        oG = oPAIR + 2 * 16;
        osz = oG + M * 16;
        dv.ensure(osz);
        hb.ensure(osz);
        PE_HIP(hipHostGetDevicePointer((void **)&hd, hb.p, 0));
        const size_t NB = (size_t)nblk;
        pRES = 0;
        pN0 = pRES + NB * 16;
        pN1 = pN0 + NB;
        pN2 = pN1 + NB;
        pC1 = pN2 + NB;
        pC2 = pC1 + NB * kCgsMax;
        pb.ensure(pC2 + NB * kCgsMax);
        G.assign((size_t)mmax * mmax, 0.0);
        dg = op->diag.p;
    }

    // Keeps the basis in parity sector sgn (0: none): the preconditioner then uses
    // the diagonal symmetrised under P, so that corrections stay in the sector bit
    // for bit (needs nd = nb^2).
    void set_sector(int sgn, i64 nb) {
        sector = sgn;
        swap_nb = nb;
        // the matrix-free sector sigma builds y(P I) from the same terms as y(I);
        // the CSR product sums the two rows in different orders
        exact = sgn == 0 || op->storage == PYCI_STORAGE_DIRECT;
        if (sgn == 0) return;
        PE_REQUIRE(nb > 0 && nb * nb == nd && nb < ((i64)1 << 31), PYCI_EINVAL, "parity sector needs nd = nb^2");
        dsym.ensure((size_t)nd);
        k_symdiag<<<g1(), 256, 0, s>>>(op->diag.p, dsym.p, nd, nb);
        PE_LAUNCH_CHECK();
        dg = dsym.p;
    }
    unsigned hnb() const { return sector != 0 ? (unsigned)swap_nb : 0u; }

    // partial[b * m + i] = block b's part of A[:, i] . x, i < m <= 16 (k_dots).
    void dots(const double *A, int m_, const double *x, double *part, const double *g0, const double *g1,
              unsigned half) {
        switch (m_) {
#define PE_DOTS(M)                                                                                  \
    case M:                                                                                         \
        k_dots<M><<<(unsigned)nblk, DT, 0, s>>>(nd, A, x, part, g0, g1, half);                    \
        break;
            PE_DOTS(1) PE_DOTS(2) PE_DOTS(3) PE_DOTS(4) PE_DOTS(5) PE_DOTS(6) PE_DOTS(7) PE_DOTS(8)
            PE_DOTS(9) PE_DOTS(10) PE_DOTS(11) PE_DOTS(12) PE_DOTS(13) PE_DOTS(14) PE_DOTS(15) PE_DOTS(16)
#undef PE_DOTS
            default: throw EngineError(PYCI_EUNSUPPORTED, "basis too large for the one-column passes");
        }
        PE_LAUNCH_CHECK();
    }

    // t -= V[:, 0:m] c (c from the partials cp), the partial squared norms to np; m <= 16.
    // With ro, the per-block partials rp are summed into *ro as well.
    void update(int m_, const double *cp, double *t, double *np, const double *g0, const double *g1,
                const double *rp = nullptr, double *ro = nullptr) {
        switch (m_) {
#define PE_UPD(M)                                                                                   \
    case M:                                                                                         \
        k_update_norm_m<M><<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, cp, t, np, g0, g1, rp, ro);     \
        break;
            PE_UPD(1) PE_UPD(2) PE_UPD(3) PE_UPD(4) PE_UPD(5) PE_UPD(6) PE_UPD(7) PE_UPD(8)
            PE_UPD(9) PE_UPD(10) PE_UPD(11) PE_UPD(12) PE_UPD(13) PE_UPD(14) PE_UPD(15) PE_UPD(16)
#undef PE_UPD
            default: throw EngineError(PYCI_EUNSUPPORTED, "basis too large for the one-column passes");
        }
        PE_LAUNCH_CHECK();
    }

    // dst (m values) = A[:, 0:m]^T x; half: A and x in the basis' parity sector
    void dots_to(const double *A, int m_, const double *x, double *dst, bool half) {
        if (m_ > 16) {
            gram(A, m_, x, 1, dst);
            return;
        }
        dots(A, m_, x, partial.p, nullptr, nullptr, half ? hnb() : 0u);
        k_reduce_partials<<<(unsigned)m_, 256, 0, s>>>(partial.p, nblk, m_, dst);
        PE_LAUNCH_CHECK();
    }

    unsigned g1() const { return (unsigned)std::min<i64>(4096, (nd + 255) / 256); }

    // out (ma x mb, row-major [i][j] on the device) = A^T B
    void gram(const double *A, int ma, const double *B, int mb, double *dst) {
        dim3 grid((unsigned)nblk, (unsigned)(((ma + 7) / 8) * ((mb + 7) / 8)));
        k_gram<<<grid, DT, 0, s>>>(nd, A, nd, ma, B, nd, mb, partial.p, nullptr, nullptr);
        PE_LAUNCH_CHECK();
        const int cnt = ma * mb;
        k_reduce_partials<<<(unsigned)cnt, 256, 0, s>>>(partial.p, nblk, cnt, dst);
        PE_LAUNCH_CHECK();
    }

    std::vector<double> gram_host(const double *A, int ma, const double *B, int mb) {
        gram(A, ma, B, mb, small.p);
        std::vector<double> h((size_t)ma * mb);
        PE_HIP(hipMemcpyAsync(h.data(), small.p, sizeof(double) * ma * mb, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        return h;  // [i][j]
    }

    // W[:, c0:c0+cnt] = H V[:, c0:c0+cnt]
    void apply(int c0, int cnt, bool fp32) {
        const double *src = V.p + (i64)c0 * nd;
        double *dst = W.p + (i64)c0 * nd;
        if (fp32) {
            row32.ensure((size_t)(nd * std::max(cnt, k)));
            out32.ensure((size_t)(nd * std::max(cnt, k)));
            k_col2row<double, float><<<g1(), 256, 0, s>>>(nd, cnt, src, nd, row32.p);
            PE_LAUNCH_CHECK();
            sigma_f32(op, cnt, row32.p, out32.p, s, sector, ws);
            k_row2col<float, double><<<g1(), 256, 0, s>>>(nd, cnt, out32.p, dst, nd);
            PE_LAUNCH_CHECK();
            st.sigma_calls_f32 += cnt;
        } else if (cnt == 1) {
            sigma_f64(op, 1, src, dst, s, sector, ws);
        } else {
            k_col2row<double, double><<<g1(), 256, 0, s>>>(nd, cnt, src, nd, rowbuf.p);
            PE_LAUNCH_CHECK();
            sigma_f64(op, cnt, rowbuf.p, rowout.p, s, sector, ws);
            k_row2col<double, double><<<g1(), 256, 0, s>>>(nd, cnt, rowout.p, dst, nd);
            PE_LAUNCH_CHECK();
        }
        st.sigma_calls += cnt;
    }

    // Orthogonalise column V[:, m] against V[:, 0:m] twice and normalise (start
    // vectors; host round trips). Returns false (column rejected) if it is
    // numerically dependent.
    bool orthonormalize_new(double norm0) {
        double *col = V.p + (i64)m * nd;
        for (int pass = 0; pass < 2 && m > 0; ++pass) {
            double *C = dv.p + oC1;
            gram(V.p, m, col, 1, C);
            k_gemm_update<<<g1(), 256, 0, s>>>(nd, V.p, nd, m, C, 1, -1.0, col, nd);
            PE_LAUNCH_CHECK();
        }
        std::vector<double> nn = gram_host(col, 1, col, 1);
        double nrm = std::sqrt(std::max(nn[0], 0.0));
        if (!(nrm > 1e-10 * std::max(norm0, 1e-300)) || !(nrm > 1e-300)) return false;
        k_scale<<<g1(), 256, 0, s>>>(col, nd, 1.0 / nrm);
        PE_LAUNCH_CHECK();
        return true;
    }

    // V[:, m] = correction j (kept in the basis' parity sector), the partials of its
    // squared norm to pN0
    void put_correction(int j) {
        k_put<<<(unsigned)nblk, DT, 0, s>>>(T.p + (i64)j * nd, V.p + (i64)m * nd, nd, swap_nb,
                                            sector != 0 ? (double)sector : 0.0, pb.p + pN0);
        PE_LAUNCH_CHECK();
    }

    // Appends column V[:, m] (filled by put_correction) orthogonalised against
    // V[:, 0:m] (classical Gram-Schmidt, the second pass only where the first lost
    // more than half the squared norm) and normalised, all on the device; the norm
    // pair (before, after) goes to the slot's pair for the host's dependency check
    // and a dependent vector is zeroed.
    void orthonormalize_device(int slot, bool have_c1 = false) {
        double *col = V.p + (i64)m * nd;
        double *pair = hd + oPAIR + 2 * slot;
        const double *p0 = pb.p + pN0;
        double *p1 = pb.p + pN1, *p2 = pb.p + pN2;
        if (m == 0) {
            k_scale_final<<<(unsigned)nblk, DT, 0, s>>>(col, nd, p0, p0, p0, nblk, 0, pair);
        } else if (m <= 16) {
            // the column is in the basis' sector: the dot products read half the rows
            if (!have_c1) dots(V.p, m, col, pb.p + pC1, nullptr, nullptr, hnb());
            // after the fused Ritz pass, the first update also finishes its residual norm
            if (have_c1)
                update(m, pb.p + pC1, col, p1, nullptr, nullptr, pb.p + pRES, hd + oRES);
            else
                update(m, pb.p + pC1, col, p1, nullptr, nullptr);
            dots(V.p, m, col, pb.p + pC2, p0, p1, hnb());
            update(m, pb.p + pC2, col, p2, p0, p1);
            k_scale_final<<<(unsigned)nblk, DT, 0, s>>>(col, nd, p0, p1, p2, nblk, 1, pair);
        } else if (m <= kCgsMax) {
            const dim3 grid((unsigned)nblk, (unsigned)((m + 7) / 8));
            k_gram<<<grid, DT, 0, s>>>(nd, V.p, nd, m, col, nd, 1, pb.p + pC1, nullptr, nullptr);
            PE_LAUNCH_CHECK();
            k_update_norm<<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, m, pb.p + pC1, col, p1, nullptr, nullptr);
            PE_LAUNCH_CHECK();
            k_gram<<<grid, DT, 0, s>>>(nd, V.p, nd, m, col, nd, 1, pb.p + pC2, p0, p1);
            PE_LAUNCH_CHECK();
            k_update_norm<<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, m, pb.p + pC2, col, p2, p0, p1);
            PE_LAUNCH_CHECK();
            k_scale_final<<<(unsigned)nblk, DT, 0, s>>>(col, nd, p0, p1, p2, nblk, 1, pair);
        } else {
            // large bases: both passes unconditionally, coefficients reduced apart
            for (int pass = 0; pass < 2; ++pass) {
                gram(V.p, m, col, 1, dv.p + oC1);
                k_gemm_update<<<g1(), 256, 0, s>>>(nd, V.p, nd, m, dv.p + oC1, 1, -1.0, col, nd);
                PE_LAUNCH_CHECK();
            }
            k_put<<<(unsigned)nblk, DT, 0, s>>>(col, col, nd, swap_nb, 0.0, p1);
            PE_LAUNCH_CHECK();
            k_scale_final<<<(unsigned)nblk, DT, 0, s>>>(col, nd, p0, p1, p1, nblk, 0, pair);
        }
        PE_LAUNCH_CHECK();
        ++m;
    }

    // Block Gram-Schmidt of the cnt corrections in V[:, m0 .. m0 + cnt): twice
    // against V[:, 0:m0] (each pass reads the basis once for the whole block), then
    // Cholesky-QR twice within the block; a column whose squared norm falls below
    // 1e-20 of its value before orthogonalisation is dependent and dropped. The
    // kept columns end orthonormal at m0 .. m0 + kept; returns kept.
    int orthonormalize_block(int m0, int cnt) {
        double *blk = V.p + (i64)m0 * nd;
        const size_t ss = (size_t)cnt * cnt;
        gram(blk, cnt, blk, cnt, dv.p + oG);  // squared norms before (diagonal)
        PE_HIP(hipMemcpyAsync(hb.p + oG, dv.p + oG, sizeof(double) * ss, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        std::vector<double> ref(cnt);
        for (int j = 0; j < cnt; ++j) ref[(size_t)j] = hb.p[oG + (size_t)j * cnt + j];
        if (m0 > 0)
            for (int pass = 0; pass < 2; ++pass) {
                gram(V.p, m0, blk, cnt, dv.p + oG);
                k_gemm_sub_rm<<<g1(), 256, 0, s>>>(nd, V.p, m0, dv.p + oG, cnt, blk);
                PE_LAUNCH_CHECK();
            }
        for (int round = 0; round < 2 && cnt > 0; ++round) {
            gram(blk, cnt, blk, cnt, dv.p + oG);
            PE_HIP(hipMemcpyAsync(hb.p + oG, dv.p + oG, sizeof(double) * (size_t)cnt * cnt, hipMemcpyDeviceToHost,
                                  s));
            PE_HIP(hipStreamSynchronize(s));
            const double *S = hb.p + oG;
            // Cholesky of S with dependent columns dropped: R (upper, over kept columns)
            std::vector<int> keep;
            std::vector<double> R((size_t)cnt * cnt, 0.0);
            for (int j = 0; j < cnt; ++j) {
                double d = S[(size_t)j * cnt + j];
                for (size_t a = 0; a < keep.size(); ++a) {
                    const int i = keep[a];
                    double x = S[(size_t)i * cnt + j];
                    for (size_t b = 0; b < a; ++b) x -= R[(size_t)keep[b] * cnt + i] * R[(size_t)keep[b] * cnt + j];
                    x /= R[(size_t)i * cnt + i];
                    R[(size_t)i * cnt + j] = x;
                    d -= x * x;
                }
                if (d > 1e-20 * std::max(ref[(size_t)j], 1e-300) && d > 0.0) {
                    R[(size_t)j * cnt + j] = std::sqrt(d);
                    keep.push_back(j);
                } else {
                    for (int i = 0; i < cnt; ++i) R[(size_t)i * cnt + j] = 0.0;
                }
            }
            // M (cnt x kept, column-major) = inverse of R restricted to the kept columns
            const int kept = (int)keep.size();
            std::vector<double> M((size_t)cnt * std::max(kept, 1), 0.0);
            for (int a = 0; a < kept; ++a) {
                // column a of R_kept^{-1}: back substitution
                std::vector<double> z((size_t)kept, 0.0);
                z[(size_t)a] = 1.0 / R[(size_t)keep[(size_t)a] * cnt + keep[(size_t)a]];
                for (int b = a - 1; b >= 0; --b) {
                    double x = 0.0;
                    for (int c = b + 1; c <= a; ++c)
                        x += R[(size_t)keep[(size_t)b] * cnt + keep[(size_t)c]] * z[(size_t)c];
                    z[(size_t)b] = -x / R[(size_t)keep[(size_t)b] * cnt + keep[(size_t)b]];
                }
                for (int b = 0; b <= a; ++b) M[(size_t)a * cnt + keep[(size_t)b]] = z[(size_t)b];
            }
            std::vector<double> nref((size_t)kept);
            for (int a = 0; a < kept; ++a) nref[(size_t)a] = ref[(size_t)keep[(size_t)a]];
            ref.swap(nref);
            if (kept == 0) {
                cnt = 0;
                break;
            }
            std::memcpy(hb.p + oYK, M.data(), sizeof(double) * M.size());
            PE_HIP(hipMemcpyAsync(dv.p + oYK, hb.p + oYK, sizeof(double) * M.size(), hipMemcpyHostToDevice, s));
            k_gemm_set<<<g1(), 256, 0, s>>>(nd, blk, nd, cnt, dv.p + oYK, kept, rowbuf.p, nd);
            PE_LAUNCH_CHECK();
            PE_HIP(hipMemcpyAsync(blk, rowbuf.p, sizeof(double) * nd * kept, hipMemcpyDeviceToDevice, s));
            cnt = kept;
        }
        PE_HIP(hipStreamSynchronize(s));  // the host-side M must outlive its copy
        for (int j = 0; j < cnt; ++j) {
            hb.p[oPAIR + 2 * j] = 1.0;
            hb.p[oPAIR + 2 * j + 1] = 1.0;
        }
        return cnt;
    }

    // Waits for the results of this iteration (residual norms; with new columns their
    // G block and norm pairs), which the kernels write straight into hb.
    void readback() { PE_HIP(hipStreamSynchronize(s)); }

    // Fills G for the cnt new columns m0..m0+cnt-1 from the read-back block and
    // removes the columns that were zeroed as dependent. Returns the number kept.
    int accept_G(int m0, int cnt) {
        const double *h = hb.p + oG, *nn = hb.p + oPAIR;
        std::vector<int> keep;
        for (int j = 0; j < cnt; ++j) {
            const double n0 = std::sqrt(std::max(nn[2 * j], 0.0)), n1 = std::sqrt(std::max(nn[2 * j + 1], 0.0));
            if (n1 > 1e-10 * std::max(n0, 1e-300) && n1 > 1e-300) keep.push_back(j);
        }
        // compact kept columns to m0, m0+1, ... (device copies; zeroed columns dropped)
        int dst = m0;
        for (int j : keep) {
            if (m0 + j != dst) {
                PE_HIP(hipMemcpyAsync(V.p + (i64)dst * nd, V.p + (i64)(m0 + j) * nd, sizeof(double) * nd,
                                      hipMemcpyDeviceToDevice, s));
                PE_HIP(hipMemcpyAsync(W.p + (i64)dst * nd, W.p + (i64)(m0 + j) * nd, sizeof(double) * nd,
                                      hipMemcpyDeviceToDevice, s));
            }
            ++dst;
        }
        // G rows/columns: old basis rows i < m0 and the kept new ones
        std::vector<int> rows;
        for (int i = 0; i < m0; ++i) rows.push_back(i);
        for (int j : keep) rows.push_back(m0 + j);
        for (int a = 0; a < (int)keep.size(); ++a) {
            const int j = keep[(size_t)a];
            for (int r = 0; r < (int)rows.size(); ++r) {
                const double x = h[(size_t)rows[(size_t)r] * cnt + j];
                G[(size_t)r * mmax + m0 + a] = x;
                G[(size_t)(m0 + a) * mmax + r] = x;
            }
        }
        for (int a = 0; a < (int)keep.size(); ++a)
            for (int b = a + 1; b < (int)keep.size(); ++b) {
                double x = 0.5 * (G[(size_t)(m0 + a) * mmax + m0 + b] + G[(size_t)(m0 + b) * mmax + m0 + a]);
                G[(size_t)(m0 + a) * mmax + m0 + b] = G[(size_t)(m0 + b) * mmax + m0 + a] = x;
            }
        m = m0 + (int)keep.size();
        return (int)keep.size();
    }

    void update_G(int c0, int cnt) {
        std::vector<double> h = gram_host(V.p, c0 + cnt, W.p + (i64)c0 * nd, cnt);  // [(c0+cnt)][cnt]
        for (int i = 0; i < c0 + cnt; ++i)
            for (int j = 0; j < cnt; ++j) {
                double x = h[(size_t)i * cnt + j];
                G[(size_t)i * mmax + c0 + j] = x;
                G[(size_t)(c0 + j) * mmax + i] = x;
            }
        // symmetrise the new diagonal block
        for (int a = 0; a < cnt; ++a)
            for (int b = a + 1; b < cnt; ++b) {
                double x = 0.5 * (G[(size_t)(c0 + a) * mmax + c0 + b] + G[(size_t)(c0 + b) * mmax + c0 + a]);
                G[(size_t)(c0 + a) * mmax + c0 + b] = G[(size_t)(c0 + b) * mmax + c0 + a] = x;
            }
    }

    // Residuals of the Ritz pairs (y: m x m eigenvectors of G, row-major; the first
    // k columns are used), with the Ritz vectors and images written when write_xw.
    void ritz(const std::vector<double> &y, const std::vector<double> &theta, bool write_xw) {
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < m; ++i) hb.p[oY + (size_t)j * m + i] = y[(size_t)i * m + j];
        for (int j = 0; j < k; ++j) hb.p[oTH + j] = theta[(size_t)j];
        PE_HIP(hipMemcpyAsync(dv.p + oY, hb.p + oY, sizeof(double) * (oTH + 16 - oY), hipMemcpyHostToDevice, s));
        y_dev = true;
        k_ritz<<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, W.p, nd, m, dv.p + oY, dv.p + oTH, k, dg,
                                            write_xw ? 1 : 0, X.p, WX.p, T.p, pb.p + pRES);
        PE_LAUNCH_CHECK();
        k_reduce_partials<<<(unsigned)k, 256, 0, s>>>(pb.p + pRES, nblk, k, hd + oRES);
        PE_LAUNCH_CHECK();
    }

    // The speculative one-root pass (k = 1, m < mmax, m <= 16): the correction goes
    // straight to V[:, m] with its squared norm (pN0) and its first Gram-Schmidt
    // coefficients (pC1), see k_ritz1.
    void ritz_spec(const std::vector<double> &y, const std::vector<double> &theta) {
        RitzCoef cf{};
        for (int i = 0; i < m; ++i) cf.y[i] = hb.p[oY + i] = y[(size_t)i * m];
        cf.theta = hb.p[oTH] = theta[0];
        y_dev = false;  // in hb only (finish_x uploads them when it needs them)
        double *t = V.p + (i64)m * nd;
        switch (m) {
#define PE_R1(M)                                                                                               \
    case M:                                                                                                    \
        k_ritz1<M><<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, W.p, cf, dg, t, pb.p + pRES, pb.p + pN0,           \
                                                 pb.p + pC1);                                                  \
        break;
            PE_R1(1) PE_R1(2) PE_R1(3) PE_R1(4) PE_R1(5) PE_R1(6) PE_R1(7) PE_R1(8)
            PE_R1(9) PE_R1(10) PE_R1(11) PE_R1(12) PE_R1(13) PE_R1(14) PE_R1(15) PE_R1(16)
#undef PE_R1
            default: throw EngineError(PYCI_EUNSUPPORTED, "basis too large for the one-column passes");
        }
        PE_LAUNCH_CHECK();  // its residual norm is summed by the first Gram-Schmidt update
    }

    // Ritz vectors X, WX of the pairs whose residuals were just read: recomputed
    // from the coefficients still on the device, or, after a restart in the same
    // iteration, the leading columns of the restarted basis.
    void finish_x(bool restarted) {
        if (restarted) {
            PE_HIP(hipMemcpyAsync(X.p, V.p, sizeof(double) * nd * k, hipMemcpyDeviceToDevice, s));
            PE_HIP(hipMemcpyAsync(WX.p, W.p, sizeof(double) * nd * k, hipMemcpyDeviceToDevice, s));
            return;
        }
        if (!y_dev) {
            PE_HIP(hipMemcpyAsync(dv.p + oY, hb.p + oY, sizeof(double) * (oTH + 16 - oY), hipMemcpyHostToDevice,
                                  s));
            y_dev = true;
        }
        k_ritz<<<(unsigned)nblk, DT, 0, s>>>(nd, V.p, W.p, nd, mx, dv.p + oY, dv.p + oTH, k, dg, 1, X.p,
                                            WX.p, T.p, pb.p + pRES);
        PE_LAUNCH_CHECK();
    }
    int mx = 0;  // basis size of the coefficients on the device (for finish_x)

    // V[:, 0:cnt] = src[:, 0:cnt], in a sector projected onto it. The basis stays in
    // its sector to the last bit by construction (every pass treats a row and its
    // partner alike); the projection at each restart bounds what rounding could
    // leave outside it, which the sector sigma does not see and the preconditioner
    // would amplify, to one restart cycle.
    void put_basis(const double *src, int cnt) {
        if (sector == 0) {
            PE_HIP(hipMemcpyAsync(V.p, src, sizeof(double) * nd * cnt, hipMemcpyDeviceToDevice, s));
            return;
        }
        for (int j = 0; j < cnt; ++j) {
            k_parity<<<g1(), 256, 0, s>>>(src + (i64)j * nd, V.p + (i64)j * nd, nd, swap_nb, (double)sector);
            PE_LAUNCH_CHECK();
        }
    }

    // Thick restart on the lowest kk Ritz vectors: they are orthonormal and G
    // becomes diag(w).
    void restart(int kk, const std::vector<double> &w, const std::vector<double> &y) {
        if (kk > k) {
            RV.ensure((size_t)(nd * 2 * k));
            RW.ensure((size_t)(nd * 2 * k));
            for (int j = 0; j < kk; ++j)
                for (int i = 0; i < m; ++i) hb.p[oYK + (size_t)j * m + i] = y[(size_t)i * m + j];
            PE_HIP(hipMemcpyAsync(dv.p + oYK, hb.p + oYK, sizeof(double) * (size_t)m * kk, hipMemcpyHostToDevice, s));
            k_gemm_set<<<g1(), 256, 0, s>>>(nd, V.p, nd, m, dv.p + oYK, kk, RV.p, nd);
            PE_LAUNCH_CHECK();
            k_gemm_set<<<g1(), 256, 0, s>>>(nd, W.p, nd, m, dv.p + oYK, kk, RW.p, nd);
            PE_LAUNCH_CHECK();
            put_basis(RV.p, kk);
            PE_HIP(hipMemcpyAsync(W.p, RW.p, sizeof(double) * nd * kk, hipMemcpyDeviceToDevice, s));
        } else {
            // the Ritz vectors of this iteration are the whole restart
            k_gemm_set<<<g1(), 256, 0, s>>>(nd, V.p, nd, m, dv.p + oY, k, X.p, nd);
            PE_LAUNCH_CHECK();
            k_gemm_set<<<g1(), 256, 0, s>>>(nd, W.p, nd, m, dv.p + oY, k, WX.p, nd);
            PE_LAUNCH_CHECK();
            put_basis(X.p, k);
            PE_HIP(hipMemcpyAsync(W.p, WX.p, sizeof(double) * nd * k, hipMemcpyDeviceToDevice, s));
        }
        std::fill(G.begin(), G.end(), 0.0);
        for (int j = 0; j < kk; ++j) G[(size_t)j * mmax + j] = w[(size_t)j];
        m = kk;
    }

    // Runs Davidson until every root meets tol (or maxiter iterations). With
    // fp32 set it stops early once residuals reach the FP32 floor (switch_tol).
    // Returns true when converged to tol; X, WX then hold the Ritz pairs.
    // One root: the correction of an iteration is orthogonalised, multiplied and
    // projected before its residual is known, so that an iteration needs a single
    // host round trip (the last iteration's correction is discarded).
    // Early stop of the second parity sector's run (link set): once the first
    // sector's run has converged, the run ends, reporting convergence, as soon as it
    // has taken at least as many iterations as the first and each of its Ritz pairs
    // has either converged (in the FP64 phase) or a Ritz value minus residual norm
    // above the first sector's highest root, so that no unconverged pair of this
    // sector can be among the k lowest roots of the two at the resolution of this
    // search. stopped_early records that this happened; an error in the first run
    // ends this one too.
    const SectorLink *link = nullptr;
    SectorLink *pub = nullptr;  // the first sector's run publishes its progress here
    bool stopped_early = false;
    // The second sector's run does not start iteration i + 1 before the first has
    // completed i: iterations past the first run's count can only be spent while
    // the first run is still going, and are wasted if its root wins.
    void pace(i64 iters) const {
        if (!link) return;
        while (link->state.load(std::memory_order_acquire) == 0 &&
               link->done.load(std::memory_order_acquire) < iters)
            std::this_thread::yield();
    }
    bool stop_now(i64 iters, const std::vector<double> &theta, const std::vector<double> &res, double tol,
                  bool fp32) const {
        if (!link) return false;
        const int st_a = link->state.load(std::memory_order_acquire);
        if (st_a == 3) throw EngineError(PYCI_EHIP, "the other parity sector's run failed");
        if (st_a != 1 || iters < link->iters) return false;
        const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
        for (int j = 0; j < k; ++j) {
            const bool above = theta[(size_t)j] - res[(size_t)j] > link->theta;
            // a root of this sector below the first's is kept only with its FP64 value
            const bool conv = !fp32 && res[(size_t)j] <= tol * std::max(std::fabs(theta[(size_t)j]), eps23);
            if (!above && !conv) return false;
        }
        return true;
    }

    bool run(double tol, i64 maxiter, bool fp32, double switch_tol, std::vector<double> &theta, i64 &iters) {
        const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
        std::vector<double> best_res(k, 1e300);
        int stall = 0;
        const bool spec = (k == 1);
        for (;;) {
            pace(iters);
            std::vector<double> gm((size_t)m * m), w, y;
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) gm[(size_t)i * m + j] = G[(size_t)i * mmax + j];
            jacobi_eig(m, gm, w, y);
            theta.assign(w.begin(), w.begin() + k);
            // one root with room for the next column: the correction comes out of the
            // Ritz pass already in its basis slot (in a sector only where the sigma keeps
            // the sector to the last bit: there is no projection on this path)
            const bool fused = spec && exact && m < nd && m + 1 <= mmax && m <= 16;
            if (fused)
                ritz_spec(y, theta);
            else
                ritz(y, theta, false);
            mx = m;
            bool restarted = false;
            int m0 = m, nspec = 0;
            if (spec && m < nd) {
                if (!fused && m + 1 > mmax) {
                    restart(std::max(k, std::min({m, 2 * k, mmax - 1})), w, y);
                    restarted = true;
                }
                m0 = m;
                if (!fused) put_correction(0);
                orthonormalize_device(0, fused);
                apply(m0, 1, fp32);
                dots_to(V.p, m0 + 1, W.p + (i64)m0 * nd, hd + oG, sector != 0 && exact);
                nspec = 1;
            }
            readback();
            std::vector<double> res(k, 0.0);
            bool all = true, floor = true;
            for (int j = 0; j < k; ++j) {
                res[j] = std::sqrt(std::max(hb.p[oRES + j], 0.0));
                if (j < 64) st.residual[j] = res[j];
                const double lim = tol * std::max(std::fabs(theta[j]), eps23);
                if (!(res[j] <= lim)) all = false;
                if (!(res[j] <= std::max(lim, switch_tol * std::max(std::fabs(theta[j]), 1.0)))) floor = false;
            }
            ++iters;
            ++st.iterations;
            if (pub) pub->done.store(iters, std::memory_order_release);
            auto leave = [&](bool ok) {
                m = m0;  // a speculative column is dropped
                finish_x(restarted);
                return ok;
            };
            if (all) return leave(true);
            if (stop_now(iters, theta, res, tol, fp32)) {
                stopped_early = true;
                return leave(true);
            }
            if (fp32 && floor) return leave(false);
            if (m0 >= nd) {
                // the basis spans the whole space: the Ritz pairs are exact to rounding
                return leave(true);
            }
            // stagnation guard: residuals at the rounding floor stop improving
            bool improved = false;
            for (int j = 0; j < k; ++j)
                if (res[j] < 0.9 * best_res[j]) {
                    best_res[j] = res[j];
                    improved = true;
                }
            stall = improved ? 0 : stall + 1;
            // stagnation: the FP32 phase hands over after 8 flat iterations once every
            // residual is within 1e-4 |theta| (near its rounding floor; a plateau far
            // above it is the normal early course of the iteration: FullCI nbasis 12
            // plateaued at |r| ~ 1 Eh, where handing over cost the FP32 phase 28 of its
            // 44 iterations), or after 40 anywhere; in FP64 a residual at the rounding
            // floor is accepted after 25, and only 300 flat iterations count as failure
            // (slow roots can plateau for dozens of iterations before they converge)
            if (fp32) {
                bool near_floor = true;
                for (int j = 0; j < k; ++j)
                    if (!(res[j] <= 1e-4 * std::max(std::fabs(theta[j]), 1.0))) near_floor = false;
                if ((stall >= 8 && near_floor) || stall >= 40) return leave(false);
            }
            if (!fp32 && stall >= 25) {
                bool tiny = true;
                for (int j = 0; j < k; ++j)
                    if (!(res[j] <= 1e-7 * std::max(std::fabs(theta[j]), 1.0))) tiny = false;
                if (tiny) return leave(true);
                if (stall >= 300) return leave(false);
            }
            if (iters >= maxiter) return leave(false);
            if (spec) {
                if (nspec) accept_G(m0, 1);
            } else {
                // unconverged roots contribute corrections
                std::vector<int> add;
                for (int j = 0; j < k; ++j)
                    if (!(res[j] <= tol * std::max(std::fabs(theta[j]), eps23))) add.push_back(j);
                if (m + (int)add.size() > mmax) {
                    restart(std::max(k, std::min({m, 2 * k, mmax - (int)add.size()})), w, y);
                    restarted = true;
                }
                m0 = m;
                const int room = (int)std::min<i64>({(i64)add.size(), (i64)(mmax - m), nd - m});
                if (room == 1) {
                    put_correction(add[0]);
                    orthonormalize_device(0);
                    apply(m0, 1, fp32);
                    dots_to(V.p, m, W.p + (i64)m0 * nd, hd + oG, sector != 0 && exact);
                    readback();
                    accept_G(m0, 1);
                } else if (room > 1) {
                    // block Gram-Schmidt: the basis is read once per pass for all corrections
                    for (int a = 0; a < room; ++a) {
                        put_correction(add[(size_t)a]);
                        ++m;
                    }
                    m = m0;
                    const int kept = orthonormalize_block(m0, room);
                    m = m0 + kept;
                    if (kept > 0) {
                        apply(m0, kept, fp32);
                        gram(V.p, m, W.p + (i64)m0 * nd, kept, hd + oG);
                        readback();
                        accept_G(m0, kept);
                    }
                }
            }
            if (m == m0) {
                // no usable direction: accept if the residuals are tiny, else fail
                bool tiny = true;
                for (int j = 0; j < k; ++j)
                    if (!(res[j] <= 1e-7 * std::max(std::fabs(theta[j]), 1.0))) tiny = false;
                // the Ritz pairs of this iteration (the coefficients are still on the device)
                finish_x(restarted);
                return tiny && !fp32;
            }
        }
    }
};

// One Davidson run for the k lowest roots (of a parity sector of P when sector
// != 0). Returns converged flag and theta (k) ascending; the Ritz vectors of the
// run stay on the device in kept.X (nd x k). stream / ws: where the run's kernels
// go and the matrix-free sigma's workspace (nullptr: the context's); kept: the
// working memory set (the context keeps one per sector stream); link: see
// Davidson::link.
static bool solve_run(pyci_op *op, int k, const double *c0, i64 mm, i64 maxit, double tol, int precision, int sector,
                      std::vector<double> &theta, DavBufs &kept, pyci_solve_stats &st,
                      const SectorLink *link = nullptr, hipStream_t stream = nullptr,
                      DevBuf<unsigned char> *ws = nullptr, bool *stopped_early = nullptr,
                      SectorLink *pub = nullptr) {
    const i64 nd = op->nrow;
    hipStream_t s = stream ? stream : op->ctx->stream;
    const i64 nb = op->swap_nb;
    auto partner = [&](i64 J) { return (J % nb) * nb + J / nb; };
    Davidson D(op, k, (int)mm, s, ws, &kept);
    D.set_sector(sector, nb);
    D.link = link;
    D.pub = pub;
    kept.hdiag.ensure((size_t)nd);
    const double *hd = kept.hdiag.p;
    PE_HIP(hipMemcpyAsync(kept.hdiag.p, op->diag.p, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    // start block: unit vectors on the lowest diagonal entries (in a sector: the
    // sector's combination of a determinant and its partner), each with a small
    // fixed perturbation; for several roots twice as many vectors as roots (at most
    // 16, leaving room for k corrections)
    const int kb0 = (k == 1) ? 1 : (int)std::min<i64>(nd, std::max<i64>(k, std::min<i64>({(i64)2 * k, 16, mm - k})));
    const i64 want = std::min<i64>(nd, 4 * (i64)kb0 + 64);
    // the `want` lowest diagonal entries in ascending order (ties by index), one pass
    // with a bounded heap
    auto lower = [&](i64 a, i64 b) { return hd[a] < hd[b] || (hd[a] == hd[b] && a < b); };
    std::vector<i64> idx;
    idx.reserve((size_t)want + 1);
    for (i64 r = 0; r < nd; ++r) {
        if ((i64)idx.size() < want) {
            idx.push_back(r);
            std::push_heap(idx.begin(), idx.end(), lower);
        } else if (lower(r, idx.front())) {
            std::pop_heap(idx.begin(), idx.end(), lower);
            idx.back() = r;
            std::push_heap(idx.begin(), idx.end(), lower);
        }
    }
    std::sort_heap(idx.begin(), idx.end(), lower);
    std::vector<double> init;
    std::vector<i64> taken;  // determinants already in a start vector of this sector
    auto is_taken = [&](i64 d) { return std::find(taken.begin(), taken.end(), d) != taken.end(); };
    size_t cursor = 0;
    D.m = 0;
    // the caller's start vector (its part in this sector), unless that part vanishes
    bool use_c0 = c0 != nullptr;
    if (use_c0 && sector != 0) {
        double all = 0.0, part = 0.0;
        for (i64 r = 0; r < nd; ++r) {
            const double x = 0.5 * (c0[r] + sector * c0[partner(r)]);
            all += c0[r] * c0[r];
            part += x * x;
        }
        use_c0 = part > 1e-16 * all && part > 0.0;
    }
    for (int j = 0; D.m < kb0 && j < kb0 + 64; ++j) {
        bool have = false;
        double *col = D.V.p + (i64)D.m * nd;
        if (j == 0 && use_c0) {
            init.assign(c0, c0 + nd);
            if (sector != 0)
                for (i64 r = 0; r < nd; ++r) init[(size_t)r] = 0.5 * (c0[r] + sector * c0[partner(r)]);
            PE_HIP(hipMemcpyAsync(col, init.data(), sizeof(double) * nd, hipMemcpyHostToDevice, s));
            have = true;
        } else {
            // perturbation of total norm ~1e-3: seeds every direction without moving
            // the start away from the dominant determinant (a perturbation of norm
            // ~0.6 leaves a start whose Rayleigh quotient is far above the lowest
            // diagonal element and can converge to an excited root); built on the device
            const u64 seed = 0x243f6a8885a308d3ULL + (u64)j;
            const double amp = 2e-3 * std::sqrt(3.0 / (double)nd);
            for (; cursor < (size_t)want && !have; ++cursor) {
                const i64 d0 = idx[cursor];
                if (sector == 0) {
                    k_start<<<D.g1(), 256, 0, s>>>(col, nd, seed, amp, 1, 0.0, d0, 1.0, -1, 0.0);
                    have = true;
                    continue;
                }
                const i64 d1 = partner(d0);
                if (is_taken(d0) || (d1 == d0 && sector < 0)) continue;
                taken.push_back(d0);
                taken.push_back(d1);
                if (d1 == d0)
                    k_start<<<D.g1(), 256, 0, s>>>(col, nd, seed, amp, nb, (double)sector, d0, 1.0, -1, 0.0);
                else
                    k_start<<<D.g1(), 256, 0, s>>>(col, nd, seed, amp, nb, (double)sector, d0, std::sqrt(0.5), d1,
                                                   sector * std::sqrt(0.5));
                have = true;
            }
            if (have) PE_LAUNCH_CHECK();
        }
        if (!have) continue;
        std::vector<double> n0 = D.gram_host(D.V.p + (i64)D.m * nd, 1, D.V.p + (i64)D.m * nd, 1);
        if (std::sqrt(std::max(n0[0], 0.0)) > 1e-12 && D.orthonormalize_new(std::sqrt(std::max(n0[0], 0.0)))) ++D.m;
    }
    PE_REQUIRE(D.m >= k, PYCI_ENOCONV, "did not converge");
    const bool mixed = (precision == PYCI_PREC_MIXED);
    D.apply(0, D.m, mixed);
    D.update_G(0, D.m);
    i64 iters = 0;
    bool ok;
    if (mixed) {
        // the FP32 phase runs to 2e-6 |theta| (an FP32 sigma is good to ~1e-7 relative,
        // so this is well above its floor; a stall of 8 iterations also hands over)
        D.run(tol, maxit, true, 2e-6, theta, iters);
        if (D.stopped_early) {
            // the sector search ended early: no FP64 phase is needed
            ok = true;
        } else {
            // FP64 residual-correction phase: restart on the FP32-phase Ritz vectors
            // with exact FP64 images.
            PE_HIP(hipMemcpyAsync(D.V.p, D.X.p, sizeof(double) * nd * k, hipMemcpyDeviceToDevice, s));
            D.m = 0;
            for (int j = 0; j < k; ++j) {
                if (j != D.m)
                    PE_HIP(hipMemcpyAsync(D.V.p + (i64)D.m * nd, D.V.p + (i64)j * nd, sizeof(double) * nd,
                                          hipMemcpyDeviceToDevice, s));
                if (D.orthonormalize_new(1.0)) ++D.m;
            }
            PE_REQUIRE(D.m == k, PYCI_ENOCONV, "did not converge");
            std::fill(D.G.begin(), D.G.end(), 0.0);
            D.apply(0, D.m, false);
            D.update_G(0, D.m);
            ok = D.run(tol, maxit, false, 0.0, theta, iters);
        }
    } else {
        ok = D.run(tol, maxit, false, 0.0, theta, iters);
    }
    st = D.st;
    if (stopped_early) *stopped_early = D.stopped_early;
    PE_HIP(hipStreamSynchronize(s));
    return ok;
}

void davidson_solve(pyci_op *op, i64 n, const double *c0, i64 ncv, i64 maxiter, double tol, int precision,
                    double *evals, double *evecs, pyci_solve_stats *stats) {
    const i64 nd = op->nrow;
    const int k = (int)n;
    PE_REQUIRE(k <= 16, PYCI_EUNSUPPORTED, "solve supports at most 16 roots per call");
    // default subspace max(14, 2n+1, 6n): thick restarts keep 2n Ritz vectors, so a
    // single root converges about as fast as with the reference's Lanczos default of
    // 20 while each iteration streams fewer basis vectors (FullCI nbasis 12, one root:
    // 131 / 125 / 125 / 135 ms with 10 / 14 / 20 / 30); blocks get 6n (fewer restarts)
    i64 mm = (ncv > 0) ? ncv : std::min<i64>(nd, std::max<i64>({2 * n + 1, 14, 6 * n}));
    mm = std::max<i64>(mm, 2 * n + 1);
    mm = std::min<i64>(mm, nd);
    mm = std::max<i64>(mm, n);
    const i64 maxit = (maxiter > 0) ? maxiter : std::max<i64>(1000, std::min<i64>(n * nd * 10, 100000));
    if (precision == PYCI_PREC_MIXED) ensure_vals32(op);
    // The alpha <-> beta exchange P commutes with H and with the diagonal
    // preconditioner, so a Davidson basis never leaves the parity sectors its start
    // vectors occupy, and roots of a sector that is not among the k lowest Ritz pairs
    // are never targeted (a random-start Lanczos, as the reference's, has no such
    // blind spot). Where P is known (full FullCI space, nup == ndn) both sectors are
    // solved and the k lowest roots of the two are returned.
    const i64 nb = op->swap_nb;
    const i64 odd_dim = (nb > 0) ? (nd - nb) / 2 : 0;
    const bool sectors = nb > 0 && odd_dim >= 2 * mm && nd - odd_dim >= 2 * mm;
    std::vector<double> theta_a, theta_b;
    // the context's working memory sets: [0] for the first (or only) run, [1] for
    // the second sector's run
    pyci_ctx *cx = op->ctx;
    for (auto &d : cx->dav)
        if (!d) d.reset(new DavBufs);
    pyci_solve_stats st_a{}, st_b{};
    std::vector<std::pair<double, std::pair<int, int>>> roots;  // (theta, (run, index))
    bool ok_a, ok_b = true;
    if (sectors) {
        // The two sectors run concurrently: the second on the context's second stream
        // and workspace from a host thread of its own, so that each run's kernels
        // fill the GPU while the other run's host does its Rayleigh-Ritz step. The
        // second sector's run is searched for as many iterations as the first needed
        // and then only as long as one of its Ritz pairs could still be among the k
        // lowest roots (SectorLink, Davidson::stop_now).
        pyci_ctx *ctx = op->ctx;
        if (!ctx->stream2) PE_HIP(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
        // the first sector's work so far (the diagonal, the FP32 values) is visible to the second stream
        PE_HIP(hipStreamSynchronize(ctx->stream));
        SectorLink link;
        bool early = false;
        std::exception_ptr err_b;
        std::thread tb([&] {
            try {
                PE_HIP(hipSetDevice(ctx->device));
                ok_b = solve_run(op, k, c0, mm, maxit, tol, precision, -1, theta_b, *cx->dav[1], st_b, &link,
                                 ctx->stream2, &ctx->mf_ws2, &early);
            } catch (...) {
                err_b = std::current_exception();
            }
        });
        try {
            ok_a = solve_run(op, k, c0, mm, maxit, tol, precision, +1, theta_a, *cx->dav[0], st_a, nullptr, nullptr,
                             nullptr, nullptr, &link);
            link.theta = theta_a[(size_t)k - 1];
            link.iters = st_a.iterations;
            link.state.store(ok_a ? 1 : 2, std::memory_order_release);
        } catch (...) {
            link.state.store(3, std::memory_order_release);
            tb.join();
            throw;
        }
        tb.join();
        if (err_b) std::rethrow_exception(err_b);
        for (int j = 0; j < k; ++j) {
            roots.push_back({theta_a[(size_t)j], {0, j}});
            roots.push_back({theta_b[(size_t)j], {1, j}});
        }
        std::stable_sort(roots.begin(), roots.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        roots.resize((size_t)k);
    } else {
        ok_a = solve_run(op, k, c0, mm, maxit, tol, precision, 0, theta_a, *cx->dav[0], st_a);
        for (int j = 0; j < k; ++j) roots.push_back({theta_a[(size_t)j], {0, j}});
    }
    if (stats) {
        *stats = st_a;
        stats->iterations += st_b.iterations;
        stats->sigma_calls += st_b.sigma_calls;
        stats->sigma_calls_f32 += st_b.sigma_calls_f32;
        for (int j = 0; j < k && j < 64; ++j) {
            const auto &r = roots[(size_t)j].second;
            stats->residual[j] = (r.first == 0 ? st_a : st_b).residual[r.second];
        }
    }
    PE_REQUIRE(ok_a && ok_b, PYCI_ENOCONV, "did not converge");
    // reference order: descending for n > 1 (sparseop.cpp:134-144)
    // the selected Ritz vectors straight from the runs' device buffers
    for (int j = 0; j < k; ++j) {
        const auto &r = roots[(size_t)((k > 1) ? (k - 1 - j) : 0)];
        evals[j] = r.first + op->ecore;
        const double *x = cx->dav[r.second.first]->X.p + (i64)r.second.second * nd;
        PE_HIP(hipMemcpyAsync(evecs + (i64)j * nd, x, sizeof(double) * nd, hipMemcpyDeviceToHost, cx->stream));
    }
    PE_HIP(hipStreamSynchronize(cx->stream));
}

}  // namespace pe
