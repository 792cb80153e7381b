// Sigma (Y = H X) over the stored full-pattern CSR: the GPU counterpart of the
// reference's Eigen SpMV in perform_op / perform_op_symm (sparseop.cpp:96-112).
// The symmetric operator is stored with both triangles, so one row-parallel
// streaming pass produces (L + L^T - D) x without scatter atomics.
//
// Layout: X is ncol x NV and Y is nrow x NV, row-major, so the NV values a column
// index points at share one cache line. One 64-lane wavefront per row; each lane
// streams column indices (4 B) and values (8 B FP64 or 4 B FP32) with a stride of
// 64 entries, so every load instruction of the wavefront is fully coalesced.
// Rows are handed to blocks in contiguous per-XCD chunks (see xcd_row_block).
#include "internal.hpp"

namespace pe {

constexpr int SPMV_WAVES = 4;  // wavefronts (rows) per 256-thread block
constexpr int NUM_XCD = 8;

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8).
// Remap so that XCD x receives a contiguous range of row blocks: neighbouring
// rows of a CI matrix share most of their column targets, so each XCD's L2 then
// serves a compact slice of X.
__device__ inline i64 xcd_row_block(i64 b, i64 nblocks) {
    const i64 per = (nblocks + NUM_XCD - 1) / NUM_XCD;
    const i64 x = b % NUM_XCD, k = b / NUM_XCD;
    const i64 r = x * per + k;
    return r;
}

// Row-major X with NV values per column; the NV values of one column are read as
// 16-byte pieces (NV * sizeof(XT) / 16 of them), and every lane issues the
// column/value loads of U entries (non-temporal: the CSR is streamed once per
// sigma) before their U * NV gathers, so each wavefront keeps U * 64 entries in flight.
template <typename XT>
struct Vec16;
template <>
struct Vec16<double> {
    typedef double2 type;
    static constexpr int n = 2;
};
template <>
struct Vec16<float> {
    typedef float4 type;
    static constexpr int n = 4;
};

// waves_per_eu(1, 4) lets the compiler hold all U * NQ pieces in registers (at
// FullCI nbasis 12, 8 FP64 vectors: 11.9 ms against 16.8 ms with the default budget).
template <int NV, int U, typename VT, typename XT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_spmm(i64 nrow, i64 nblocks, const i64 *__restrict__ rowptr, const int32_t *__restrict__ cols,
       const VT *__restrict__ vals, const XT *__restrict__ X, XT *__restrict__ Y) {
    typedef typename Vec16<XT>::type V16;
    constexpr int PER = Vec16<XT>::n, NQ = NV / PER;  // XT values per piece, pieces per column
    static_assert(NV % PER == 0, "NV must fill whole 16-byte pieces");
    const int lane = threadIdx.x & 63;
    const i64 blk = xcd_row_block(blockIdx.x, nblocks);
    const i64 row = blk * SPMV_WAVES + (threadIdx.x >> 6);
    if (blk >= nblocks || row >= nrow) return;
    const i64 beg = rowptr[row], end = rowptr[row + 1];
    XT acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = (XT)0;
    const V16 *__restrict__ X16 = reinterpret_cast<const V16 *>(X);
    i64 p = beg + lane;
    for (; p + (U - 1) * 64 < end; p += U * 64) {
        int c[U];
        VT v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = __builtin_nontemporal_load(&cols[p + u * 64]);
            v[u] = __builtin_nontemporal_load(&vals[p + u * 64]);
        }
        V16 x[U][NQ];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQ; ++q) x[u][q] = X16[(i64)c[u] * NQ + q];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const XT w = (XT)v[u];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const XT *e = reinterpret_cast<const XT *>(&x[u][q]);
#pragma unroll
                for (int t = 0; t < PER; ++t) acc[q * PER + t] += w * e[t];
            }
        }
    }
    for (; p < end; p += 64) {
        const int c = __builtin_nontemporal_load(&cols[p]);
        const XT w = (XT)__builtin_nontemporal_load(&vals[p]);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const V16 xv = X16[(i64)c * NQ + q];
            const XT *e = reinterpret_cast<const XT *>(&xv);
#pragma unroll
            for (int t = 0; t < PER; ++t) acc[q * PER + t] += w * e[t];
        }
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        XT v = acc[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        acc[j] = v;
    }
    if (lane < NV) {
        XT out = acc[0];
#pragma unroll
        for (int j = 1; j < NV; ++j)
            if (lane == j) out = acc[j];
        Y[row * NV + lane] = out;
    }
}

// Single vector: every lane issues U column/value loads (non-temporal: the CSR is
// streamed once per sigma and must not evict X from L2) before their U dependent
// X gathers, so each wavefront keeps U*64 entries in flight.
template <int U, typename VT, typename XT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_spmv(i64 nrow, i64 nblocks, const i64 *__restrict__ rowptr, const int32_t *__restrict__ cols,
       const VT *__restrict__ vals, const XT *__restrict__ X, XT *__restrict__ Y) {
    const int lane = threadIdx.x & 63;
    const i64 blk = xcd_row_block(blockIdx.x, nblocks);
    const i64 row = blk * SPMV_WAVES + (threadIdx.x >> 6);
    if (blk >= nblocks || row >= nrow) return;
    const i64 beg = rowptr[row], end = rowptr[row + 1];
    XT acc = (XT)0;
    i64 p = beg + lane;
    for (; p + (U - 1) * 64 < end; p += U * 64) {
        int c[U];
        VT v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = __builtin_nontemporal_load(&cols[p + u * 64]);
            v[u] = __builtin_nontemporal_load(&vals[p + u * 64]);
        }
        XT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = X[c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (XT)v[u] * x[u];
    }
    for (; p < end; p += 64) acc += (XT)__builtin_nontemporal_load(&vals[p]) * X[__builtin_nontemporal_load(&cols[p])];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) Y[row] = acc;
}

// Generic nvec: loops over vector chunks of 8.
template <typename VT, typename XT>
__global__ void __launch_bounds__(256) k_spmm_any(i64 nrow, i64 nblocks, int nv, const i64 *__restrict__ rowptr,
                                                  const int32_t *__restrict__ cols, const VT *__restrict__ vals,
                                                  const XT *__restrict__ X, XT *__restrict__ Y) {
    const int lane = threadIdx.x & 63;
    const i64 blk = xcd_row_block(blockIdx.x, nblocks);
    const i64 row = blk * SPMV_WAVES + (threadIdx.x >> 6);
    if (blk >= nblocks || row >= nrow) return;
    const i64 beg = rowptr[row], end = rowptr[row + 1];
    for (int j0 = 0; j0 < nv; j0 += 8) {
        const int cnt = min(8, nv - j0);
        XT acc[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = (XT)0;
        for (i64 p = beg + lane; p < end; p += 64) {
            const int c = cols[p];
            const XT v = (XT)vals[p];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < cnt) acc[j] += v * X[(i64)c * nv + j0 + j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            XT v = acc[j];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            acc[j] = v;
        }
        if (lane < cnt) {
            XT out = acc[0];
#pragma unroll
            for (int j = 1; j < 8; ++j)
                if (lane == j) out = acc[j];
            Y[row * nv + j0 + lane] = out;
        }
    }
}

template <typename VT, typename XT>
static void spmm(const pyci_op *op, const VT *vals, int nvec, const XT *X, XT *Y, hipStream_t s) {
    const i64 nrow = op->nrow;
    if (nrow <= 0) return;
    const i64 nblocks = (nrow + SPMV_WAVES - 1) / SPMV_WAVES;
    const i64 grid = ((nblocks + NUM_XCD - 1) / NUM_XCD) * NUM_XCD;
    const i64 *rp = op->rowptr.p;
    const int32_t *c = op->cols.p;
    switch (nvec) {
        case 1:
            // 16 (FP64) or 8 (FP32) entries in flight per lane: 5.7 / 4.4 ms at FullCI
            // nbasis 12 against 6.0 / 4.7 ms with 4
            if constexpr (sizeof(XT) == 8)
                k_spmv<16, VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, rp, c, vals, X, Y);
            else
                k_spmv<8, VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, rp, c, vals, X, Y);
            break;
        case 2:
            // two FP64 values are one 16-byte piece; two FP32 values are not
            if constexpr (sizeof(XT) == 8)
                k_spmm<2, 4, VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, rp, c, vals, X, Y);
            else
                k_spmm_any<VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, nvec, rp, c, vals, X, Y);
            break;
        case 4: k_spmm<4, 4, VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, rp, c, vals, X, Y); break;
        case 8: k_spmm<8, 4, VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, rp, c, vals, X, Y); break;
        default: k_spmm_any<VT, XT><<<(unsigned)grid, 256, 0, s>>>(nrow, nblocks, nvec, rp, c, vals, X, Y); break;
    }
    PE_LAUNCH_CHECK();
}

void sigma_f64(const pyci_op *op, int nvec, const double *X, double *Y, hipStream_t s, int sector,
               DevBuf<unsigned char> *ws) {
    if (op->storage == PYCI_STORAGE_DIRECT) return direct_sigma<double>(op, nvec, X, Y, s, nullptr, sector, ws);
    spmm<double, double>(op, op->vals.p, nvec, X, Y, s);
}

void sigma_f32(const pyci_op *op, int nvec, const float *X, float *Y, hipStream_t s, int sector,
               DevBuf<unsigned char> *ws) {
    if (op->storage == PYCI_STORAGE_DIRECT) return direct_sigma<float>(op, nvec, X, Y, s, nullptr, sector, ws);
    spmm<float, float>(op, op->vals32.p, nvec, X, Y, s);
}

__global__ void k_to_f32(const double *a, float *b, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

void ensure_vals32(pyci_op *op) {
    if (op->storage == PYCI_STORAGE_DIRECT) return;  // the matrix-free operator has FP32 integrals of its own
    if (op->vals32.p && op->vals32.n >= (size_t)op->nnz) return;
    op->vals32.alloc((size_t)op->nnz);
    if (op->nnz > 0) {
        k_to_f32<<<4096, 256, 0, op->ctx->stream>>>(op->vals.p, op->vals32.p, op->nnz);
        PE_LAUNCH_CHECK();
    }
}

}  // namespace pe

namespace pe {

__global__ void k_fill_random(double *x, i64 n, u64 seed) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        x[i] = (double)(mix64(seed + (u64)i) >> 11) * 0x1p-53 - 0.5;
}

// Times nrep launches of the sigma kernel with HIP events on the context stream.
// Algorithmic bytes of one launch: every stored entry's column index and value,
// the row pointers, one read of X and one write of Y. For the matrix-free
// operator: one read of X, one write of Y, and one write and one read each of
// the pair-indexed intermediates D and E (npair values per determinant and vector).
void op_bench_sigma(pyci_op *op, i64 nvec, i64 nrep, int fp32, float *avg_ms, double *bytes) {
    hipStream_t s = op->ctx->stream;
    const i64 nr = op->nrow, nc = op->ncol;
    PE_REQUIRE(nvec >= 1 && nrep >= 1, PYCI_EINVAL, "nvec and nrep must be >= 1");
    const int vb = fp32 ? 4 : 8;
    if (fp32) ensure_vals32(op);
    DevBuf<double> x((size_t)(nc * nvec)), y((size_t)(nr * nvec));
    DevBuf<float> x32, y32;
    k_fill_random<<<1024, 256, 0, s>>>(x.p, nc * nvec, 12345);
    PE_LAUNCH_CHECK();
    if (fp32) {
        x32.alloc((size_t)(nc * nvec));
        y32.alloc((size_t)(nr * nvec));
        k_to_f32<<<1024, 256, 0, s>>>(x.p, x32.p, nc * nvec);
        PE_LAUNCH_CHECK();
    }
    auto launch = [&] {
        if (fp32)
            sigma_f32(op, (int)nvec, x32.p, y32.p, s);
        else
            sigma_f64(op, (int)nvec, x.p, y.p, s);
    };
    launch();  // warm-up
    PE_HIP(hipEventRecord(op->ctx->ev0, s));
    for (i64 r = 0; r < nrep; ++r) launch();
    PE_HIP(hipEventRecord(op->ctx->ev1, s));
    PE_HIP(hipEventSynchronize(op->ctx->ev1));
    float ms = 0.f;
    PE_HIP(hipEventElapsedTime(&ms, op->ctx->ev0, op->ctx->ev1));
    *avg_ms = ms / (float)nrep;
    if (op->storage == PYCI_STORAGE_DIRECT)
        *bytes = (double)vb * (double)nvec * (double)nr * (2.0 + 4.0 * (double)op->direct->npair);
    else
        *bytes = (double)op->nnz * (4.0 + vb) + 8.0 * (double)(nr + 1) + (double)vb * (double)nvec * (double)(nc + nr);
}

}  // namespace pe
