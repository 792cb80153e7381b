// General CSR builder: one 64-lane wavefront per determinant row.
//
// Reference: SparseOp::update -> add_row -> sort_row (sparseop.cpp:186-502).
// Every lane takes candidate excitations e = lane, lane+64, ... from a flat
// index over the reference's excitation classes (FullCI: 1-0, 1-1, 2-0, 0-1,
// 0-2; DOCI: pair hops; GenCI: singles, doubles), forms the excited
// determinant, looks up its column through the device index, and keeps it when
// column < ncol. Matrix elements use the reference's formulas and summation
// order, so values are bit-identical to the serial code.
//
// Pass 1 counts entries per row (all stored entries, and those with col <= row
// for the lower-triangle export), an exclusive scan turns counts into row
// pointers, pass 2 writes entries (wave-level ballot compaction) and the
// diagonal, and a per-row bitonic sort in LDS orders each row by column.
#include "internal.hpp"

#pragma clang fp contract(off)

namespace pe {

constexpr int ROWS_PER_BLOCK = 4;  // 4 wavefronts x 64 lanes

__device__ inline int nth_bit(u64 x, int t) {
    for (int s = 0; s < t && x; ++s) x &= x - 1;
    return x ? __builtin_ctzll(x) : 0;
}

// t = k(k-1)/2 + i with i < k
__device__ inline void tri_decode(int t, int &i, int &k) {
    int kk = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
    while (kk * (kk - 1) / 2 > t) --kk;
    while ((kk + 1) * kk / 2 <= t) ++kk;
    k = kk;
    i = t - kk * (kk - 1) / 2;
}

#define GIX(p, q, r, s) ((((p) * n + (q)) * n + (r)) * n + (s))

template <int KIND, bool FILL>
__global__ void __launch_bounds__(256) k_rows(WfnView w, HamView H, i64 row0, i64 row1, i64 ncol, i64 *cnt_full,
                                              i64 *cnt_low, int *maxlen, const i64 *rowptr, int32_t *cols,
                                              double *vals, double *diag) {
    const int lane = threadIdx.x & 63;
    const i64 row = row0 + (i64)blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
    if (row >= row1) return;  // uniform per wavefront; no block barriers below
    const int n = H.n;
    const double *one = H.one, *two = H.two;
    const u64 m = full_mask(n);
    const u64 up = w.dets[(i64)w.nkey * row];
    const u64 dn = (KIND == PYCI_WFN_FULLCI) ? w.dets[2 * row + 1] : 0ULL;
    const int nou = __popcll(up), nvu = n - nou;
    const int nod = __popcll(dn), nvd = n - nod;
    // lane q holds the q-th occupied/virtual orbital of each string, packed
    const int packed = nth_bit(up, lane) | (nth_bit(~up & m, lane) << 8) | (nth_bit(dn, lane) << 16) |
                       (nth_bit(~dn & m, lane) << 24);
    // candidate classes
    int c0, c1, c2, c3, c4;  // cumulative class boundaries
    const int nsu = nou * nvu, nsd = nod * nvd;
    const int npo_u = nou * (nou - 1) / 2, npv_u = nvu * (nvu - 1) / 2;
    const int npo_d = nod * (nod - 1) / 2, npv_d = nvd * (nvd - 1) / 2;
    if (KIND == PYCI_WFN_FULLCI) {
        c0 = nsu;                  // 1-0
        c1 = c0 + nsu * nsd;       // 1-1
        c2 = c1 + npo_u * npv_u;   // 2-0
        c3 = c2 + nsd;             // 0-1
        c4 = c3 + npo_d * npv_d;   // 0-2
    } else if (KIND == PYCI_WFN_DOCI) {
        c0 = nsu;
        c1 = c2 = c3 = c4 = c0;
    } else {
        c0 = nsu;                  // singles
        c1 = c0;
        c2 = c1 + npo_u * npv_u;   // doubles
        c3 = c4 = c2;
    }
    const int ncand = c4;
    i64 out_base = 0;
    if (FILL) out_base = rowptr[row - row0];
    int written = 0, cfull = 0, clow = 0;
    for (int e0 = 0; e0 < ncand; e0 += 64) {
        const int e = e0 + lane;
        const bool valid = e < ncand;
        int cls = 0, i = 0, a = 0, k = 0, l = 0;
        if (valid) {
            if (e < c0) {
                cls = 0;
                i = e / nvu;
                a = e - i * nvu;
            } else if (e < c1) {
                cls = 1;
                int t = e - c0, ia = t / nsd, kl = t - ia * nsd;
                i = ia / nvu;
                a = ia - i * nvu;
                k = kl / nvd;
                l = kl - k * nvd;
            } else if (e < c2) {
                cls = 2;
                int t = e - c1, po = t / npv_u, pv = t - po * npv_u;
                tri_decode(po, i, k);
                tri_decode(pv, a, l);
            } else if (e < c3) {
                cls = 3;
                int t = e - c2;
                i = t / nvd;
                a = t - i * nvd;
            } else {
                cls = 4;
                int t = e - c3, po = t / npv_d, pv = t - po * npv_d;
                tri_decode(po, i, k);
                tri_decode(pv, a, l);
            }
        }
        // gather orbital numbers from the owning lanes (uniform control flow)
        const int pi = __shfl(packed, i), pa = __shfl(packed, a), pk = __shfl(packed, k), pl = __shfl(packed, l);
        i64 col = -1;
        double val = 0.0;
        if (valid) {
            const bool alpha = (cls <= 2);
            const int ii = alpha ? (pi & 255) : ((pi >> 16) & 255);
            const int jj = alpha ? ((pa >> 8) & 255) : ((pa >> 24) & 255);
            int kk, ll;
            if (cls == 1) {
                kk = (pk >> 16) & 255;
                ll = (pl >> 24) & 255;
            } else if (alpha) {
                kk = pk & 255;
                ll = (pl >> 8) & 255;
            } else {
                kk = (pk >> 16) & 255;
                ll = (pl >> 24) & 255;
            }
            u64 nu = up, nd = dn;
            if (KIND == PYCI_WFN_DOCI) {
                nu = up ^ (1ULL << ii) ^ (1ULL << jj);
            } else if (cls == 0) {
                nu = up ^ (1ULL << ii) ^ (1ULL << jj);
            } else if (cls == 1) {
                nu = up ^ (1ULL << ii) ^ (1ULL << jj);
                nd = dn ^ (1ULL << kk) ^ (1ULL << ll);
            } else if (cls == 2) {
                nu = up ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
            } else if (cls == 3) {
                nd = dn ^ (1ULL << ii) ^ (1ULL << jj);
            } else {
                nd = dn ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
            }
            col = wfn_lookup(w, nu, nd);
            if (col >= ncol) col = -1;
            if (FILL && col >= 0) {
                if (KIND == PYCI_WFN_DOCI) {
                    val = two[GIX(ii, ii, jj, jj)];  // v[k,l] = <kk|ll>
                } else if (cls == 0) {
                    double v = one[ii * n + jj];
                    for (u64 x = up; x; x &= x - 1) {
                        int q = __builtin_ctzll(x);
                        v += two[GIX(ii, q, jj, q)] - two[GIX(ii, q, q, jj)];
                    }
                    if (KIND == PYCI_WFN_FULLCI)
                        for (u64 x = dn; x; x &= x - 1) {
                            int q = __builtin_ctzll(x);
                            v += two[GIX(ii, q, jj, q)];
                        }
                    val = (double)phase1(up, ii, jj) * v;
                } else if (cls == 1) {
                    val = (double)(phase1(up, ii, jj) * phase1(dn, kk, ll)) * two[GIX(ii, kk, jj, ll)];
                } else if (cls == 2) {
                    val = (double)phase2(up, ii, kk, jj, ll) * (two[GIX(ii, kk, jj, ll)] - two[GIX(ii, kk, ll, jj)]);
                } else if (cls == 3) {
                    double v = one[ii * n + jj];
                    for (u64 x = up; x; x &= x - 1) {
                        int q = __builtin_ctzll(x);
                        v += two[GIX(ii, q, jj, q)];
                    }
                    for (u64 x = dn; x; x &= x - 1) {
                        int q = __builtin_ctzll(x);
                        v += two[GIX(ii, q, jj, q)] - two[GIX(ii, q, q, jj)];
                    }
                    val = (double)phase1(dn, ii, jj) * v;
                } else {
                    val = (double)phase2(dn, ii, kk, jj, ll) * (two[GIX(ii, kk, jj, ll)] - two[GIX(ii, kk, ll, jj)]);
                }
            }
        }
        const u64 hit = __ballot(col >= 0);
        if (FILL) {
            if (col >= 0) {
                const int pos = written + __popcll(hit & ((1ULL << lane) - 1ULL));
                cols[out_base + pos] = (int32_t)col;
                vals[out_base + pos] = val;
            }
            written += __popcll(hit);
        } else {
            cfull += __popcll(hit);
            clow += __popcll(__ballot(col >= 0 && col < row));
        }
    }
    if (lane == 0) {
        const bool has_diag = row < ncol;
        if (FILL) {
            double d;
            if (KIND == PYCI_WFN_DOCI) {
                double v1 = 0.0, v2 = 0.0;
                for (u64 x = up; x; x &= x - 1) {
                    int kk = __builtin_ctzll(x);
                    v1 += two[GIX(kk, kk, kk, kk)];
                    v2 += one[kk * (n + 1)];
                    for (u64 y = x & (x - 1); y; y &= y - 1) {
                        int ll = __builtin_ctzll(y);
                        v2 += two[GIX(kk, ll, kk, ll)] * 2 - two[GIX(kk, ll, ll, kk)];
                    }
                }
                d = v1 + v2 * 2;
            } else {
                double v2 = 0.0;
                for (u64 x = up; x; x &= x - 1) {
                    int ii = __builtin_ctzll(x);
                    v2 += one[(n + 1) * ii];
                    for (u64 y = x & (x - 1); y; y &= y - 1) {
                        int kk = __builtin_ctzll(y);
                        v2 += two[GIX(ii, kk, ii, kk)] - two[GIX(ii, kk, kk, ii)];
                    }
                    if (KIND == PYCI_WFN_FULLCI)
                        for (u64 y = dn; y; y &= y - 1) {
                            int kk = __builtin_ctzll(y);
                            v2 += two[GIX(ii, kk, ii, kk)];
                        }
                }
                if (KIND == PYCI_WFN_FULLCI)
                    for (u64 x = dn; x; x &= x - 1) {
                        int ii = __builtin_ctzll(x);
                        v2 += one[(n + 1) * ii];
                        for (u64 y = x & (x - 1); y; y &= y - 1) {
                            int kk = __builtin_ctzll(y);
                            v2 += two[GIX(ii, kk, ii, kk)] - two[GIX(ii, kk, kk, ii)];
                        }
                    }
                d = v2;
            }
            diag[row - row0] = d;
            if (has_diag) {
                cols[out_base + written] = (int32_t)row;
                vals[out_base + written] = d;
            }
        } else {
            if (has_diag) {
                ++cfull;
                ++clow;
            }
            cnt_full[row - row0] = cfull;
            cnt_low[row - row0] = clow;
            atomicMax(maxlen, cfull);
        }
    }
}

// ---------------------------------------------------------------------------
// Per-row sort by column (sort_row, sparseop.cpp:214-218). Bitonic network in
// the all-ascending form, so padding past the row end behaves as +infinity.
// ---------------------------------------------------------------------------
constexpr int SORT_CAP = 4096;
constexpr int SORT_THREADS = 256;

__global__ void __launch_bounds__(SORT_THREADS) k_sort_rows_lds(i64 nrow, const i64 *rowptr, int32_t *cols,
                                                                double *vals) {
    __shared__ int32_t sk[SORT_CAP];
    __shared__ double sv[SORT_CAP];
    const i64 row = blockIdx.x;
    if (row >= nrow) return;
    const i64 b = rowptr[row];
    const int len = (int)(rowptr[row + 1] - b);
    if (len <= 1 || len > SORT_CAP) return;  // uniform per block
    int N = 1;
    while (N < len) N <<= 1;
    for (int t = threadIdx.x; t < N; t += SORT_THREADS) {
        if (t < len) {
            sk[t] = cols[b + t];
            sv[t] = vals[b + t];
        } else {
            sk[t] = 0x7fffffff;
        }
    }
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
        const int half = size >> 1;
        for (int t = threadIdx.x; t < N / 2; t += SORT_THREADS) {
            int blk = t / half, o = t - blk * half;
            int p = blk * size + o, q = blk * size + size - 1 - o;
            if (sk[q] < sk[p]) {
                int32_t tk = sk[p];
                sk[p] = sk[q];
                sk[q] = tk;
                double tv = sv[p];
                sv[p] = sv[q];
                sv[q] = tv;
            }
        }
        __syncthreads();
        for (int stride = half >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < N / 2; t += SORT_THREADS) {
                int blk = t / stride, o = t - blk * stride;
                int p = blk * 2 * stride + o, q = p + stride;
                if (sk[q] < sk[p]) {
                    int32_t tk = sk[p];
                    sk[p] = sk[q];
                    sk[q] = tk;
                    double tv = sv[p];
                    sv[p] = sv[q];
                    sv[q] = tv;
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < len; t += SORT_THREADS) {
        cols[b + t] = sk[t];
        vals[b + t] = sv[t];
    }
}

// Rows longer than SORT_CAP: the same network directly on global memory.
__global__ void __launch_bounds__(SORT_THREADS) k_sort_rows_global(i64 nrow, const i64 *rowptr, int32_t *cols,
                                                                   double *vals) {
    const i64 row = blockIdx.x;
    if (row >= nrow) return;
    const i64 b = rowptr[row];
    const i64 len = rowptr[row + 1] - b;
    if (len <= SORT_CAP) return;
    int32_t *k = cols + b;
    double *v = vals + b;
    i64 N = 1;
    while (N < len) N <<= 1;
    for (i64 size = 2; size <= N; size <<= 1) {
        const i64 half = size >> 1;
        for (i64 t = threadIdx.x; t < N / 2; t += SORT_THREADS) {
            i64 blk = t / half, o = t - blk * half;
            i64 p = blk * size + o, q = blk * size + size - 1 - o;
            if (q < len && k[q] < k[p]) {
                int32_t tk = k[p];
                k[p] = k[q];
                k[q] = tk;
                double tv = v[p];
                v[p] = v[q];
                v[q] = tv;
            }
        }
        __syncthreads();
        for (i64 stride = half >> 1; stride > 0; stride >>= 1) {
            for (i64 t = threadIdx.x; t < N / 2; t += SORT_THREADS) {
                i64 blk = t / stride, o = t - blk * stride;
                i64 p = blk * 2 * stride + o, q = p + stride;
                if (q < len && k[q] < k[p]) {
                    int32_t tk = k[p];
                    k[p] = k[q];
                    k[q] = tk;
                    double tv = v[p];
                    v[p] = v[q];
                    v[q] = tv;
                }
            }
            __syncthreads();
        }
    }
}

template <int KIND>
static void launch_rows(bool fill, const WfnView &w, const HamView &H, i64 row0, i64 row1, i64 ncol, i64 *cf,
                        i64 *cl, int *mx, const i64 *rp, int32_t *cols, double *vals, double *diag, hipStream_t s) {
    unsigned g = grid_for(row1 - row0, ROWS_PER_BLOCK);
    if (fill)
        k_rows<KIND, true><<<g, 256, 0, s>>>(w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag);
    else
        k_rows<KIND, false><<<g, 256, 0, s>>>(w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag);
    PE_LAUNCH_CHECK();
}

static void dispatch_rows(int kind, bool fill, const WfnView &w, const HamView &H, i64 row0, i64 row1, i64 ncol,
                          i64 *cf, i64 *cl, int *mx, const i64 *rp, int32_t *cols, double *vals, double *diag,
                          hipStream_t s) {
    if (row1 <= row0) return;
    if (kind == PYCI_WFN_DOCI)
        launch_rows<PYCI_WFN_DOCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s);
    else if (kind == PYCI_WFN_FULLCI)
        launch_rows<PYCI_WFN_FULLCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s);
    else
        launch_rows<PYCI_WFN_GENCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s);
}

void build_rows(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn, i64 row0, i64 row1, i64 ncol,
                DevBuf<i64> &rowptr_seg, DevBuf<int32_t> &cols, DevBuf<double> &vals, DevBuf<i64> &lowcnt_seg,
                DevBuf<double> &diag_seg, i64 &nnz_seg, i64 &low_total) {
    pyci_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const i64 nr = row1 - row0;
    WfnView w = view_of(wfn);
    HamView H = view_of(ham);
    rowptr_seg.alloc((size_t)(nr + 1));
    lowcnt_seg.alloc((size_t)(nr + 1));
    diag_seg.alloc((size_t)(nr > 0 ? nr : 1));
    DevBuf<int> mx(1);
    PE_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), s));
    PE_HIP(hipMemsetAsync(rowptr_seg.p, 0, sizeof(i64) * (nr + 1), s));
    PE_HIP(hipMemsetAsync(lowcnt_seg.p, 0, sizeof(i64) * (nr + 1), s));
    dispatch_rows(op->kind, false, w, H, row0, row1, ncol, rowptr_seg.p, lowcnt_seg.p, mx.p, nullptr, nullptr,
                  nullptr, nullptr, s);
    // lower-triangle total (reference size) on the host from the low counts
    {
        DevBuf<i64> tmp((size_t)(nr + 1));
        PE_HIP(hipMemcpyAsync(tmp.p, lowcnt_seg.p, sizeof(i64) * (nr + 1), hipMemcpyDeviceToDevice, s));
        low_total = exclusive_scan(ctx, tmp.p, nr + 1);
    }
    nnz_seg = exclusive_scan(ctx, rowptr_seg.p, nr + 1);
    int hmax = 0;
    PE_HIP(hipMemcpyAsync(&hmax, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    PE_REQUIRE(nnz_seg < (i64)1 << 40, PYCI_ENOMEM, "sparse operator too large");
    cols.alloc((size_t)nnz_seg);
    vals.alloc((size_t)nnz_seg);
    dispatch_rows(op->kind, true, w, H, row0, row1, ncol, nullptr, nullptr, nullptr, rowptr_seg.p, cols.p, vals.p,
                  diag_seg.p, s);
    if (nr > 0) {
        k_sort_rows_lds<<<(unsigned)nr, SORT_THREADS, 0, s>>>(nr, rowptr_seg.p, cols.p, vals.p);
        PE_LAUNCH_CHECK();
        if (hmax > SORT_CAP) {
            k_sort_rows_global<<<(unsigned)nr, SORT_THREADS, 0, s>>>(nr, rowptr_seg.p, cols.p, vals.p);
            PE_LAUNCH_CHECK();
        }
    }
}

// ---------------------------------------------------------------------------
// Export in the reference format: the first lowcnt[row] entries of each sorted
// row (symmetric) or all entries, as int64 columns and float64 values.
// ---------------------------------------------------------------------------
__global__ void k_export(i64 nrow, const i64 *rowptr, const i64 *outptr, const int32_t *cols, const double *vals,
                         i64 *ocols, double *ovals) {
    const int lane = threadIdx.x & 63;
    const i64 row = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrow) return;
    const i64 src = rowptr[row], dst = outptr[row], len = outptr[row + 1] - dst;
    for (i64 t = lane; t < len; t += 64) {
        ocols[dst + t] = cols[src + t];
        ovals[dst + t] = vals[src + t];
    }
}

void export_csr(pyci_op *op, double *data, i64 *indices, i64 *indptr) {
    hipStream_t s = op->ctx->stream;
    const i64 nr = op->nrow;
    DevBuf<i64> outptr((size_t)(nr + 1));
    if (op->symmetric) {
        PE_HIP(hipMemcpyAsync(outptr.p, op->lowcnt.p, sizeof(i64) * (nr + 1), hipMemcpyDeviceToDevice, s));
        exclusive_scan(op->ctx, outptr.p, nr + 1);
    } else {
        PE_HIP(hipMemcpyAsync(outptr.p, op->rowptr.p, sizeof(i64) * (nr + 1), hipMemcpyDeviceToDevice, s));
    }
    const i64 size = op->size_ref;
    DevBuf<i64> oc((size_t)(size > 0 ? size : 1));
    DevBuf<double> ov((size_t)(size > 0 ? size : 1));
    if (nr > 0) {
        k_export<<<grid_for(nr, 4), 256, 0, s>>>(nr, op->rowptr.p, outptr.p, op->cols.p, op->vals.p, oc.p, ov.p);
        PE_LAUNCH_CHECK();
    }
    if (size > 0) {
        PE_HIP(hipMemcpyAsync(data, ov.p, sizeof(double) * size, hipMemcpyDeviceToHost, s));
        PE_HIP(hipMemcpyAsync(indices, oc.p, sizeof(i64) * size, hipMemcpyDeviceToHost, s));
    }
    PE_HIP(hipMemcpyAsync(indptr, outptr.p, sizeof(i64) * (nr + 1), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

// get_element(i, j): binary search in the sorted row (sparseop.cpp:89-94).
__global__ void k_get_element(const i64 *rowptr, const int32_t *cols, const double *vals, i64 i, i64 j,
                              double *out) {
    i64 lo = rowptr[i], hi = rowptr[i + 1];
    double r = 0.0;
    while (lo < hi) {
        i64 mid = (lo + hi) >> 1;
        if ((i64)cols[mid] < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < rowptr[i + 1] && (i64)cols[lo] == j) r = vals[lo];
    *out = r;
}

double get_element(pyci_op *op, i64 i, i64 j) {
    hipStream_t s = op->ctx->stream;
    DevBuf<double> d(1);
    k_get_element<<<1, 1, 0, s>>>(op->rowptr.p, op->cols.p, op->vals.p, i, j, d.p);
    PE_LAUNCH_CHECK();
    double h = 0.0;
    PE_HIP(hipMemcpyAsync(&h, d.p, sizeof(double), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace pe
