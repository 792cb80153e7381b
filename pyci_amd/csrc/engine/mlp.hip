// Learned selector: a 2-hidden-layer MLP (occupation bits -> 256 -> 256 -> 1,
// ReLU) predicting log10|c| of a determinant, trained and evaluated with BF16
// MFMA kernels (v_mfma_f32_32x32x16_bf16, FP32 accumulation) on gfx950.
//
// No counterpart in the reference; it ranks the heat-bath candidates of
// add_hci (hci.cpp:252-279) so that the selection keeps the candidates with
// the largest predicted amplitudes.
//
// Every matrix product is one kernel, C[M][N] = A[M][K] . Bt[N][K]^T, with both
// operands K-contiguous so that each lane's MFMA fragment (8 consecutive k) is
// one 16-byte load. Activations are kept in both layouts where a later product
// needs them with the sample index as k:
//   forward   z1 = x  W1   : A = x   [R][64],   Bt = w1t [256][64]   -> h1, h1t
//             z2 = h1 W2   : A = h1  [R][256],  Bt = w2t [256][256]  -> h2t, y partials
//   backward  dh1 = dz2 W2^T: A = dz2 [R][256], Bt = w2  [256][256]  -> dz1t (ReLU mask)
//             dW2 = h1^T dz2: A = h1t [256][R], Bt = dz2t [256][R]   (split over samples)
//             dW1 = x^T dz1 : A = xt  [128][R], Bt = dz1t [256][R]   (rows >= 64 are zero)
// All reductions run in a fixed order, so results do not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace pe {

constexpr int HID = 256;
constexpr int FPAD = 64;
constexpr i64 OW1 = 0, OB1 = OW1 + FPAD * HID, OW2 = OB1 + HID, OB2 = OW2 + HID * HID, OW3 = OB2 + HID,
              OB3 = OW3 + HID, NPARAM = OB3 + 1;
static_assert(NPARAM == PYCI_MLP_NPARAM, "parameter layout");
constexpr i64 KCHUNK = 512;  // samples per split of the weight-gradient products

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
    // round to nearest even (finite inputs)
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline bf16x8 ld_frag(const uint16_t *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

enum { EPI_F32 = 0, EPI_L1 = 1, EPI_L2 = 2, EPI_BWD1 = 3 };

struct Epi {
    const float *bias = nullptr;      // L1, L2
    const float *w3 = nullptr;        // L2
    uint16_t *h = nullptr;            // L1: row-major [m][n], ld = N
    uint16_t *ht = nullptr;           // L1, L2, BWD1: transposed [n][m], ld = ldt
    i64 ldt = 0;
    float *ypart = nullptr;           // L2: [N/64][M]
    const uint16_t *mask_t = nullptr; // BWD1: h1t [n][m], ld = ldt
    float *out = nullptr;             // F32: [split][M][N]
};

// 256 threads = 4 waves; block tile 128 x 128, wave tile 64 x 64 (2 x 2 MFMA
// tiles of 32 x 32). Requires M % 128 == 0, N % 128 == 0, K % 16 == 0 and
// kchunk % 16 == 0 (checked by the host). Fragment layout of
// v_mfma_f32_32x32x16_bf16: lane l supplies row (l % 32) of A and column
// (l % 32) of B with k = 8 (l / 32) + e, e < 8; accumulator register r of lane l
// holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
template <int MODE>
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, i64 K, const uint16_t *__restrict__ A, i64 lda,
                                                   const uint16_t *__restrict__ Bt, i64 ldb, i64 kchunk, Epi ep) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 31, hh = lane >> 5;
    const int m0 = blockIdx.x * 128 + (wave >> 1) * 64;
    const int n0 = blockIdx.y * 128 + (wave & 1) * 64;
    const i64 kb = (i64)blockIdx.z * kchunk;
    const i64 ke = (kb + kchunk < K) ? kb + kchunk : K;
    const uint16_t *a0 = A + (i64)(m0 + r) * lda + 8 * hh;
    const uint16_t *a1 = a0 + 32 * lda;
    const uint16_t *b0 = Bt + (i64)(n0 + r) * ldb + 8 * hh;
    const uint16_t *b1 = b0 + 32 * ldb;
    f32x16 c[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) c[i][j] = f32x16{};
#pragma unroll 2
    for (i64 k = kb; k < ke; k += 16) {
        const bf16x8 fa0 = ld_frag(a0 + k), fa1 = ld_frag(a1 + k);
        const bf16x8 fb0 = ld_frag(b0 + k), fb1 = ld_frag(b1 + k);
        c[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa0, fb0, c[0][0], 0, 0, 0);
        c[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa0, fb1, c[0][1], 0, 0, 0);
        c[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa1, fb0, c[1][0], 0, 0, 0);
        c[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa1, fb1, c[1][1], 0, 0, 0);
    }
    if (MODE == EPI_F32) {
        float *o = ep.out + (i64)blockIdx.z * M * N;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int m = m0 + 32 * i + 8 * (q >> 2) + 4 * hh + (q & 3), n = n0 + 32 * j + r;
                    o[(i64)m * N + n] = c[i][j][q];
                }
        return;
    }
    if (MODE == EPI_L1 || MODE == EPI_BWD1) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = n0 + 32 * j + r;
                const float b = (MODE == EPI_L1) ? ep.bias[n] : 0.0f;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int mb = m0 + 32 * i + 8 * g + 4 * hh;  // 4 consecutive rows mb..mb+3
                    uint16_t v[4];
                    if (MODE == EPI_BWD1) {
                        const uint2 mk = *reinterpret_cast<const uint2 *>(ep.mask_t + (i64)n * ep.ldt + mb);
                        const uint16_t mm[4] = {(uint16_t)(mk.x & 0xffff), (uint16_t)(mk.x >> 16),
                                                (uint16_t)(mk.y & 0xffff), (uint16_t)(mk.y >> 16)};
                        for (int e = 0; e < 4; ++e) v[e] = (mm[e] != 0) ? f2bf(c[i][j][4 * g + e]) : (uint16_t)0;
                    } else {
                        for (int e = 0; e < 4; ++e) {
                            const float z = c[i][j][4 * g + e] + b;
                            v[e] = f2bf(z > 0.0f ? z : 0.0f);
                            ep.h[(i64)(mb + e) * N + n] = v[e];
                        }
                    }
                    if (ep.ht) {
                        uint2 pk;
                        pk.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
                        pk.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
                        *reinterpret_cast<uint2 *>(ep.ht + (i64)n * ep.ldt + mb) = pk;
                    }
                }
            }
        return;
    }
    // EPI_L2: h2 = bf16(relu(z2 + b2)); partial y over this wave's 64 columns
    float ysum[2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 16; ++q) ysum[i][q] = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + 32 * j + r;
        const float b = ep.bias[n], w = ep.w3[n];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int mb = m0 + 32 * i + 8 * g + 4 * hh;
                uint16_t v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float z = c[i][j][4 * g + e] + b;
                    v[e] = f2bf(z > 0.0f ? z : 0.0f);
                    ysum[i][4 * g + e] += bf2f(v[e]) * w;
                }
                if (ep.ht) {
                    uint2 pk;
                    pk.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
                    pk.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
                    *reinterpret_cast<uint2 *>(ep.ht + (i64)n * ep.ldt + mb) = pk;
                }
            }
    }
    float *yp = ep.ypart + (i64)(n0 >> 6) * M;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            float s = ysum[i][q];
            for (int o = 16; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (r == 0) yp[m0 + 32 * i + 8 * (q >> 2) + 4 * hh + (q & 3)] = s;
        }
}

// Features of R samples: x [R][64] and xt [128][R] (bf16 0/1). Sample s < n is
// determinant sel(s): s + first when elig == nullptr, otherwise a counter-based
// draw from the eligible list (with its target copied); rows s >= n are zero.
__global__ void k_features(int kind, int nbasis, const u64 *__restrict__ dets, int nkey, i64 first, i64 n, i64 R,
                           const i64 *__restrict__ elig, i64 nelig, u64 seed, const float *__restrict__ tall,
                           float *__restrict__ tgt, uint16_t *__restrict__ x, uint16_t *__restrict__ xt) {
    const i64 s = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;  // feature group of 8
    if (s >= R) return;
    u64 a = 0, b = 0;
    if (s < n) {
        i64 d = s + first;
        if (elig) {
            d = elig[(i64)(mix64(seed + 0x9e3779b97f4a7c15ULL * (u64)(s + 1)) % (u64)nelig)];
            if (g == 0) tgt[s] = tall[d];
        }
        a = dets[(i64)nkey * d];
        if (nkey == 2) b = dets[2 * d + 1];
    } else if (elig && g == 0) {
        tgt[s] = 0.0f;
    }
    uint16_t v[8];
    for (int e = 0; e < 8; ++e) {
        const int f = 8 * g + e;
        int bit = 0;
        if (kind == PYCI_WFN_FULLCI) {
            if (f < nbasis) bit = (int)((a >> f) & 1ULL);
            else if (f < 2 * nbasis) bit = (int)((b >> (f - nbasis)) & 1ULL);
        } else if (f < nbasis) {
            bit = (int)((a >> f) & 1ULL);
        }
        v[e] = bit ? (uint16_t)0x3f80 : (uint16_t)0;
        xt[(i64)f * R + s] = v[e];
    }
    uint4 pk;
    pk.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    pk.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    pk.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
    pk.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
    *reinterpret_cast<uint4 *>(x + s * FPAD + 8 * g) = pk;
}

// y = ((p0 + p1) + p2) + p3 + b3; with targets also dy = 2 (y - t) / n (0 for
// rows >= n) and the loss sum / n in out[0], db3 = sum dy in out[1]. One block.
__global__ void __launch_bounds__(1024) k_y_loss(const float *ypart, i64 R, i64 n, const float *params,
                                                 const float *tgt, float *y, float *dy, float *out) {
    const float b3 = params[OB3];
    float ls = 0.0f, ds = 0.0f;
    for (i64 m = threadIdx.x; m < R; m += 1024) {
        const float v = ((ypart[m] + ypart[R + m]) + ypart[2 * R + m]) + ypart[3 * R + m] + b3;
        y[m] = v;
        if (tgt) {
            float d = 0.0f;
            if (m < n) {
                const float e = v - tgt[m];
                ls += e * e;
                d = 2.0f * e / (float)n;
            }
            dy[m] = d;
            ds += d;
        }
    }
    if (!tgt) return;
    __shared__ float sl[1024], sd[1024];
    sl[threadIdx.x] = ls;
    sd[threadIdx.x] = ds;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sl[threadIdx.x] += sl[threadIdx.x + o];
            sd[threadIdx.x] += sd[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sl[0] / (float)n;
        out[1] = sd[0];
    }
}

// dz2 = bf16(dy[m] w3[f] [h2 > 0]) into dz2 [R][256] and dz2t [256][R], via a
// 64 x 64 LDS transpose. Block (64, 4); grid (R / 64, 256 / 64).
__global__ void __launch_bounds__(256) k_dz2(const uint16_t *__restrict__ h2t, const float *__restrict__ dy,
                                             const float *__restrict__ w3, i64 R, uint16_t *__restrict__ dz2,
                                             uint16_t *__restrict__ dz2t) {
    __shared__ uint16_t tile[64][65];
    const i64 mb = (i64)blockIdx.x * 64;
    const int fb = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const float d = dy[mb + tx];
    for (int fy = ty; fy < 64; fy += 4) {
        const int f = fb + fy;
        const i64 o = (i64)f * R + mb + tx;
        const uint16_t hv = h2t[o];
        const uint16_t v = (hv != 0) ? f2bf(d * w3[f]) : (uint16_t)0;
        dz2t[o] = v;
        tile[fy][tx] = v;
    }
    __syncthreads();
    for (int my = ty; my < 64; my += 4) dz2[(mb + my) * HID + fb + tx] = tile[tx][my];
}

// out[row] = sum_m X[row][m] (* w[m]); one 256-thread block per row, fixed order.
__global__ void __launch_bounds__(256) k_rowsum(const uint16_t *__restrict__ X, i64 ld, i64 ncols,
                                                const float *__restrict__ w, float *__restrict__ out) {
    const uint16_t *row = X + (i64)blockIdx.x * ld;
    float s = 0.0f;
    for (i64 m = threadIdx.x; m < ncols; m += 256) s += w ? bf2f(row[m]) * w[m] : bf2f(row[m]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// dst[i] = sum_z part[z * stride + i], z ascending.
__global__ void k_sum_splits(const float *part, int nsplit, i64 stride, i64 count, float *dst) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < nsplit; ++z) s += part[(i64)z * stride + i];
    dst[i] = s;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, i64 n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// bf16 operand copies of the weights.
__global__ void k_refresh(const float *p, uint16_t *w1t, uint16_t *w2t, uint16_t *w2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < HID * HID) {
        const int f1 = i / HID, f2 = i % HID;
        const uint16_t v = f2bf(p[OW2 + i]);
        w2[i] = v;
        w2t[f2 * HID + f1] = v;
    }
    if (i < FPAD * HID) {
        const int in = i / HID, f1 = i % HID;
        w1t[f1 * FPAD + in] = f2bf(p[OW1 + i]);
    }
}

__global__ void k_targets(const double *c, i64 n, double cmin, float *t, int *ok) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = fabs(c[i]);
    t[i] = (float)log10(a > cmin ? a : cmin);
    ok[i] = a >= cmin ? 1 : 0;
}

__global__ void k_int_flags(const int *f, i64 n, i64 *dst) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) dst[i] = (i < n) ? (i64)f[i] : 0;
}

__global__ void k_compact_index(const int *f, const i64 *pos, i64 n, i64 *out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && f[i]) out[pos[i]] = i;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static i64 round128(i64 n) { return (n + 127) / 128 * 128; }

static void ensure_workspace(pyci_mlp *m, i64 R) {
    if (R <= m->cap) return;
    hipStream_t s = m->ctx->stream;
    PE_HIP(hipStreamSynchronize(s));
    m->x.alloc((size_t)R * FPAD);
    m->xt.alloc((size_t)128 * R);
    PE_HIP(hipMemsetAsync(m->xt.p, 0, sizeof(uint16_t) * 128 * R, s));  // rows >= 64 stay zero
    m->h1.alloc((size_t)R * HID);
    m->h1t.alloc((size_t)R * HID);
    m->h2t.alloc((size_t)R * HID);
    m->dz2.alloc((size_t)R * HID);
    m->dz2t.alloc((size_t)R * HID);
    m->dz1t.alloc((size_t)R * HID);
    m->y.alloc((size_t)R);
    m->ypart.alloc((size_t)4 * R);
    m->dy.alloc((size_t)R);
    m->tgt.alloc((size_t)R);
    const i64 nsplit = (R + KCHUNK - 1) / KCHUNK;
    m->wpart.alloc((size_t)nsplit * HID * HID);
    m->red.alloc(8);
    m->cap = R;
}

template <int MODE>
static void gemm(hipStream_t s, i64 M, i64 N, i64 K, const uint16_t *A, i64 lda, const uint16_t *Bt, i64 ldb,
                 i64 kchunk, const Epi &ep) {
    PE_REQUIRE(M % 128 == 0 && N % 128 == 0 && K % 16 == 0 && kchunk % 16 == 0 && lda % 8 == 0 && ldb % 8 == 0,
               PYCI_EINVAL, "mlp product shape");
    PE_REQUIRE(M < (1LL << 31) && N < (1LL << 31), PYCI_EINVAL, "mlp product too large");
    const unsigned nz = (unsigned)((K + kchunk - 1) / kchunk);
    dim3 grid((unsigned)(M / 128), (unsigned)(N / 128), nz);
    k_gemm_bf16<MODE><<<grid, 256, 0, s>>>((int)M, (int)N, K, A, lda, Bt, ldb, kchunk, ep);
    PE_LAUNCH_CHECK();
}

// Forward over R rows already featurised in m->x; with `train` the activations
// needed by the backward are kept (h1t, h2t).
static void forward(pyci_mlp *m, i64 R, bool train) {
    hipStream_t s = m->ctx->stream;
    const float *p = m->params.p;
    Epi e1;
    e1.bias = p + OB1;
    e1.h = m->h1.p;
    e1.ht = train ? m->h1t.p : nullptr;
    e1.ldt = R;
    gemm<EPI_L1>(s, R, HID, FPAD, m->x.p, FPAD, m->w1t.p, FPAD, FPAD, e1);
    Epi e2;
    e2.bias = p + OB2;
    e2.w3 = p + OW3;
    e2.ht = train ? m->h2t.p : nullptr;
    e2.ldt = R;
    e2.ypart = m->ypart.p;
    gemm<EPI_L2>(s, R, HID, HID, m->h1.p, HID, m->w2t.p, HID, HID, e2);
}

// Loss and gradient (into m->grads) of the R featurised rows, n of them real,
// with targets in m->tgt. Returns nothing; the loss is left in m->red[0].
static void loss_backward(pyci_mlp *m, i64 R, i64 n) {
    hipStream_t s = m->ctx->stream;
    float *g = m->grads.p;
    forward(m, R, true);
    k_y_loss<<<1, 1024, 0, s>>>(m->ypart.p, R, n, m->params.p, m->tgt.p, m->y.p, m->dy.p, m->red.p);
    PE_LAUNCH_CHECK();
    k_dz2<<<dim3((unsigned)(R / 64), HID / 64), 256, 0, s>>>(m->h2t.p, m->dy.p, m->params.p + OW3, R, m->dz2.p,
                                                            m->dz2t.p);
    PE_LAUNCH_CHECK();
    Epi eb;
    eb.ht = m->dz1t.p;
    eb.ldt = R;
    eb.mask_t = m->h1t.p;
    gemm<EPI_BWD1>(s, R, HID, HID, m->dz2.p, HID, m->w2.p, HID, HID, eb);
    const int nsplit = (int)((R + KCHUNK - 1) / KCHUNK);
    Epi ew;
    ew.out = m->wpart.p;
    gemm<EPI_F32>(s, HID, HID, R, m->h1t.p, R, m->dz2t.p, R, KCHUNK, ew);
    k_sum_splits<<<grid_for(HID * HID, 256), 256, 0, s>>>(m->wpart.p, nsplit, HID * HID, HID * HID, g + OW2);
    PE_LAUNCH_CHECK();
    gemm<EPI_F32>(s, 128, HID, R, m->xt.p, R, m->dz1t.p, R, KCHUNK, ew);
    k_sum_splits<<<grid_for(FPAD * HID, 256), 256, 0, s>>>(m->wpart.p, nsplit, 128 * HID, FPAD * HID, g + OW1);
    PE_LAUNCH_CHECK();
    k_rowsum<<<HID, 256, 0, s>>>(m->dz1t.p, R, R, nullptr, g + OB1);
    PE_LAUNCH_CHECK();
    k_rowsum<<<HID, 256, 0, s>>>(m->dz2t.p, R, R, nullptr, g + OB2);
    PE_LAUNCH_CHECK();
    k_rowsum<<<HID, 256, 0, s>>>(m->h2t.p, R, R, m->dy.p, g + OW3);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(g + OB3, m->red.p + 1, sizeof(float), hipMemcpyDeviceToDevice, s));
}

static void refresh(pyci_mlp *m) {
    k_refresh<<<HID * HID / 256, 256, 0, m->ctx->stream>>>(m->params.p, m->w1t.p, m->w2t.p, m->w2.p);
    PE_LAUNCH_CHECK();
}

static void featurise(pyci_mlp *m, int kind, i64 nbasis, int nkey, const u64 *d_dets, i64 first, i64 n, i64 R,
                      const i64 *elig, i64 nelig, u64 seed, const float *tall) {
    dim3 grid(grid_for(R, 256), FPAD / 8);
    k_features<<<grid, 256, 0, m->ctx->stream>>>(kind, (int)nbasis, d_dets, nkey, first, n, R, elig, nelig, seed,
                                                 tall, m->tgt.p, m->x.p, m->xt.p);
    PE_LAUNCH_CHECK();
}

static int nfeat_of(int kind, i64 nbasis) { return (int)(kind == PYCI_WFN_FULLCI ? 2 * nbasis : nbasis); }

static void check_features(pyci_mlp *m, int kind, i64 nbasis) {
    PE_REQUIRE(kind == PYCI_WFN_DOCI || kind == PYCI_WFN_FULLCI || kind == PYCI_WFN_GENCI, PYCI_EINVAL,
               "unknown wave function kind");
    PE_REQUIRE(nfeat_of(kind, nbasis) == m->nfeat, PYCI_EINVAL,
               "determinant features (" + std::to_string(nfeat_of(kind, nbasis)) + ") do not match the network (" +
                   std::to_string(m->nfeat) + ")");
}

void mlp_init(pyci_mlp *m, int nfeat, u64 seed) {
    PE_REQUIRE(nfeat >= 1 && nfeat <= FPAD, PYCI_EINVAL, "the network takes 1..64 input features");
    m->nfeat = nfeat;
    m->params.alloc(NPARAM);
    m->grads.alloc(NPARAM);
    m->adam_m.alloc(NPARAM);
    m->adam_v.alloc(NPARAM);
    m->w1t.alloc(HID * FPAD);
    m->w2t.alloc(HID * HID);
    m->w2.alloc(HID * HID);
    // He-uniform weights from a splitmix64 stream; zero biases
    std::vector<float> p(NPARAM, 0.0f);
    u64 st = seed;
    auto uni = [&]() {
        st += 0x9e3779b97f4a7c15ULL;
        u64 z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const double a1 = std::sqrt(6.0 / nfeat), a2 = std::sqrt(6.0 / HID), a3 = std::sqrt(3.0 / HID);
    for (int in = 0; in < nfeat; ++in)
        for (int f = 0; f < HID; ++f) p[OW1 + in * HID + f] = (float)(a1 * uni());
    for (i64 i = 0; i < HID * HID; ++i) p[OW2 + i] = (float)(a2 * uni());
    for (int f = 0; f < HID; ++f) p[OW3 + f] = (float)(a3 * uni());
    mlp_set_params(m, p.data());
}

void mlp_set_params(pyci_mlp *m, const float *host) {
    hipStream_t s = m->ctx->stream;
    PE_HIP(hipMemcpyAsync(m->params.p, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, s));
    PE_HIP(hipMemsetAsync(m->adam_m.p, 0, sizeof(float) * NPARAM, s));
    PE_HIP(hipMemsetAsync(m->adam_v.p, 0, sizeof(float) * NPARAM, s));
    m->adam_t = 0;
    refresh(m);
    PE_HIP(hipStreamSynchronize(s));
}

void mlp_get_params(pyci_mlp *m, float *host) {
    hipStream_t s = m->ctx->stream;
    PE_HIP(hipMemcpyAsync(host, m->params.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

constexpr i64 PRED_CHUNK = 65536;

// y[i] for n device determinants, chunked.
void mlp_predict_device(pyci_mlp *m, int kind, i64 nbasis, const u64 *d_dets, i64 n, float *d_y) {
    check_features(m, kind, nbasis);
    hipStream_t s = m->ctx->stream;
    const int nkey = kind == PYCI_WFN_FULLCI ? 2 : 1;
    ensure_workspace(m, std::min(round128(std::max<i64>(n, 1)), PRED_CHUNK));
    for (i64 first = 0; first < n; first += PRED_CHUNK) {
        const i64 cnt = std::min(PRED_CHUNK, n - first), R = round128(cnt);
        featurise(m, kind, nbasis, nkey, d_dets, first, cnt, R, nullptr, 0, 0, nullptr);
        forward(m, R, false);
        k_y_loss<<<1, 1024, 0, s>>>(m->ypart.p, R, cnt, m->params.p, nullptr, m->y.p, nullptr, nullptr);
        PE_LAUNCH_CHECK();
        PE_HIP(hipMemcpyAsync(d_y + first, m->y.p, sizeof(float) * cnt, hipMemcpyDeviceToDevice, s));
    }
}

void mlp_predict(pyci_mlp *m, int kind, i64 nbasis, i64 n, const u64 *dets, float *y) {
    if (n <= 0) return;
    const int nkey = kind == PYCI_WFN_FULLCI ? 2 : 1;
    hipStream_t s = m->ctx->stream;
    DevBuf<u64> dd((size_t)(n * nkey));
    DevBuf<float> dy((size_t)n);
    PE_HIP(hipMemcpyAsync(dd.p, dets, sizeof(u64) * n * nkey, hipMemcpyHostToDevice, s));
    mlp_predict_device(m, kind, nbasis, dd.p, n, dy.p);
    PE_HIP(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

float mlp_loss_grad(pyci_mlp *m, int kind, i64 nbasis, i64 n, const u64 *dets, const float *target, float *grad) {
    check_features(m, kind, nbasis);
    PE_REQUIRE(n >= 1, PYCI_EINVAL, "empty batch");
    const int nkey = kind == PYCI_WFN_FULLCI ? 2 : 1;
    hipStream_t s = m->ctx->stream;
    const i64 R = round128(n);
    ensure_workspace(m, R);
    DevBuf<u64> dd((size_t)(n * nkey));
    PE_HIP(hipMemcpyAsync(dd.p, dets, sizeof(u64) * n * nkey, hipMemcpyHostToDevice, s));
    PE_HIP(hipMemsetAsync(m->tgt.p, 0, sizeof(float) * R, s));
    PE_HIP(hipMemcpyAsync(m->tgt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, s));
    featurise(m, kind, nbasis, nkey, dd.p, 0, n, R, nullptr, 0, 0, nullptr);
    loss_backward(m, R, n);
    float loss = 0.0f;
    PE_HIP(hipMemcpyAsync(&loss, m->red.p, sizeof(float), hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(grad, m->grads.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    return loss;
}

float mlp_train(pyci_mlp *m, const pyci_wfn *wfn, const double *coeffs, double cmin, i64 steps, i64 batch, double lr,
                u64 seed) {
    check_features(m, wfn->kind, wfn->nbasis);
    PE_REQUIRE(steps >= 0 && batch >= 128 && batch % 128 == 0, PYCI_EINVAL,
               "batch must be a positive multiple of 128");
    PE_REQUIRE(cmin > 0.0, PYCI_EINVAL, "cmin must be positive");
    hipStream_t s = m->ctx->stream;
    const i64 nd = wfn->ndet;
    PE_REQUIRE(nd >= 1, PYCI_EINVAL, "empty wave function");
    DevBuf<double> dc((size_t)nd);
    DevBuf<float> tall((size_t)nd);
    DevBuf<int> ok((size_t)nd);
    DevBuf<i64> pos((size_t)(nd + 1));
    PE_HIP(hipMemcpyAsync(dc.p, coeffs, sizeof(double) * nd, hipMemcpyHostToDevice, s));
    k_targets<<<grid_for(nd, 256), 256, 0, s>>>(dc.p, nd, cmin, tall.p, ok.p);
    PE_LAUNCH_CHECK();
    k_int_flags<<<grid_for(nd + 1, 256), 256, 0, s>>>(ok.p, nd, pos.p);
    PE_LAUNCH_CHECK();
    const i64 nelig = exclusive_scan(m->ctx, pos.p, nd + 1);
    PE_REQUIRE(nelig >= 1, PYCI_EINVAL, "no determinant has |c| >= cmin");
    DevBuf<i64> elig((size_t)nelig);
    k_compact_index<<<grid_for(nd, 256), 256, 0, s>>>(ok.p, pos.p, nd, elig.p);
    PE_LAUNCH_CHECK();
    if (m->adam_t == 0) {
        // a fresh network starts from the mean target as its output bias
        std::vector<double> hc((size_t)nd);
        PE_HIP(hipMemcpyAsync(hc.data(), dc.p, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        i64 cnt = 0;
        for (double c : hc)
            if (std::fabs(c) >= cmin) sum += std::log10(std::fabs(c)), ++cnt;
        const float b3 = (float)(sum / (double)cnt);
        PE_HIP(hipMemcpyAsync(m->params.p + OB3, &b3, sizeof(float), hipMemcpyHostToDevice, s));
        PE_HIP(hipStreamSynchronize(s));
    }
    const i64 R = batch;
    ensure_workspace(m, R);
    const int nkey = wfn->nkey;
    const float b1 = 0.9f, b2 = 0.999f;
    for (i64 t = 0; t < steps; ++t) {
        featurise(m, wfn->kind, wfn->nbasis, nkey, wfn->dets.p, 0, R, R, elig.p, nelig,
                  seed ^ (0xd1b54a32d192ed03ULL * (u64)(m->adam_t + 1)), tall.p);
        loss_backward(m, R, R);
        m->adam_t += 1;
        const float bc1 = 1.0f - std::pow(b1, (float)m->adam_t), bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
        k_adam<<<grid_for(NPARAM, 256), 256, 0, s>>>(m->params.p, m->grads.p, m->adam_m.p, m->adam_v.p, NPARAM,
                                                     (float)lr, b1, b2, 1e-8f, bc1, bc2);
        PE_LAUNCH_CHECK();
        refresh(m);
    }
    float loss = 0.0f;
    PE_HIP(hipMemcpyAsync(&loss, m->red.p, sizeof(float), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    return loss;
}

// ---------------------------------------------------------------------------
// Selection: top-k candidates by score, ties in candidate order.
// ---------------------------------------------------------------------------
__device__ inline uint32_t order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_scores(const float *y, const double *hb, i64 n, int mode, uint32_t *key) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = y[i];
    if (mode == 1) v += (float)log10(hb[i] > 1e-300 ? hb[i] : 1e-300);
    key[i] = order_key(v);
}

__global__ void __launch_bounds__(256) k_hist(const uint32_t *key, i64 n, uint32_t prefix, uint32_t mask, int shift,
                                              unsigned long long *hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const uint32_t k = key[i];
        if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_tie_flags(const uint32_t *key, i64 n, uint32_t T, i64 *tie) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) tie[i] = (i < n && key[i] == T) ? 1 : 0;
}

__global__ void k_take_flags(const uint32_t *key, const i64 *tie_rank, i64 n, uint32_t T, i64 need, i64 *take) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    take[i] = (i < n && (key[i] > T || (key[i] == T && tie_rank[i] < need))) ? 1 : 0;
}

__global__ void k_take_write(const uint32_t *key, const i64 *tie_rank, const i64 *pos, i64 n, uint32_t T, i64 need,
                             const u64 *cand, int nkey, u64 *out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (key[i] > T || (key[i] == T && tie_rank[i] < need)) {
        const i64 o = pos[i];
        for (int w = 0; w < nkey; ++w) out[(i64)nkey * o + w] = cand[(i64)nkey * i + w];
    }
}

void wfn_index_append(pyci_wfn *w, i64 first);

i64 add_hci_mlp_device(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs_host, double eps, pyci_mlp *m,
                       double pool, i64 nadd, int mode) {
    PE_REQUIRE(pool >= 1.0, PYCI_EINVAL, "pool must be >= 1");
    PE_REQUIRE(mode == 0 || mode == 1, PYCI_EINVAL, "score_mode must be 0 or 1");
    check_features(m, wfn->kind, wfn->nbasis);
    pyci_ctx *ctx = wfn->ctx;
    hipStream_t s = ctx->stream;
    const i64 nd = wfn->ndet;
    if (nd == 0) return 0;
    DevBuf<double> dc((size_t)nd);
    PE_HIP(hipMemcpyAsync(dc.p, coeffs_host, sizeof(double) * nd, hipMemcpyHostToDevice, s));
    HciPool P;
    hci_pool_device(ham, wfn, dc.p, eps, eps / pool, P);
    const i64 n = P.n;
    const i64 k = std::min(nadd < 0 ? P.nplain : nadd, n);
    if (k <= 0) return 0;
    DevBuf<float> y((size_t)n);
    mlp_predict_device(m, wfn->kind, wfn->nbasis, P.dets.p, n, y.p);
    DevBuf<uint32_t> key((size_t)n);
    k_scores<<<grid_for(n, 256), 256, 0, s>>>(y.p, P.hb.p, n, mode, key.p);
    PE_LAUNCH_CHECK();
    // radix select of the k-th largest key, 8 bits per pass
    DevBuf<unsigned long long> hist(256);
    uint32_t prefix = 0, mask = 0;
    i64 remaining = k;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        PE_HIP(hipMemsetAsync(hist.p, 0, sizeof(unsigned long long) * 256, s));
        k_hist<<<std::min<unsigned>(grid_for(n, 256), 2048), 256, 0, s>>>(key.p, n, prefix, mask, shift, hist.p);
        PE_LAUNCH_CHECK();
        unsigned long long h[256];
        PE_HIP(hipMemcpyAsync(h, hist.p, sizeof(h), hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        int b = 255;
        for (; b > 0; --b) {
            if ((i64)h[b] >= remaining) break;
            remaining -= (i64)h[b];
        }
        prefix |= (uint32_t)b << shift;
        mask |= 255u << shift;
    }
    const uint32_t T = prefix;  // key of the k-th largest; `remaining` ties with it are taken
    DevBuf<i64> tie((size_t)(n + 1)), take((size_t)(n + 1));
    k_tie_flags<<<grid_for(n + 1, 256), 256, 0, s>>>(key.p, n, T, tie.p);
    PE_LAUNCH_CHECK();
    exclusive_scan(ctx, tie.p, n + 1);
    k_take_flags<<<grid_for(n + 1, 256), 256, 0, s>>>(key.p, tie.p, n, T, remaining, take.p);
    PE_LAUNCH_CHECK();
    const i64 ntake = exclusive_scan(ctx, take.p, n + 1);
    PE_REQUIRE(ntake == k, PYCI_EHIP, "selection count mismatch");
    const i64 need = (nd + k) * wfn->nkey;
    if ((size_t)need > wfn->dets.n) {
        DevBuf<u64> nb((size_t)(need + need / 2));
        PE_HIP(hipMemcpyAsync(nb.p, wfn->dets.p, sizeof(u64) * nd * wfn->nkey, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipStreamSynchronize(s));
        wfn->dets = std::move(nb);
    }
    k_take_write<<<grid_for(n, 256), 256, 0, s>>>(key.p, tie.p, take.p, n, T, remaining, P.dets.p, wfn->nkey,
                                                  wfn->dets.p + nd * wfn->nkey);
    PE_LAUNCH_CHECK();
    PE_HIP(hipStreamSynchronize(s));
    wfn->ndet = nd + k;
    wfn_index_append(wfn, nd);
    PE_HIP(hipStreamSynchronize(s));
    return k;
}

}  // namespace pe
