// FanCI objective and Jacobian on the GPU (fanci.cpp:129-217, overlaps of
// ap1rog.cpp:55-220 and apig.cpp:55-195).
//
// The reference evaluates the Jacobian one parameter at a time: a perform_op of
// each column of d(c_m)/d(p_k) (fanci.cpp:197-204). Here the derivative matrix is
// built on the device in the layout the CSR product reads (row-major: the nparam
// values of one determinant are contiguous) and multiplied by the rectangular
// operator in ONE SpMM, then the Jacobian is assembled in place (column-major, as
// the reference returns it).
//
// Overlaps are permanents of small matrices of the geminal coefficients
// (Ryser's formula, the reference's loop order): one thread per determinant for
// the overlaps, one thread per (determinant, parameter) for the derivatives.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace pe {

constexpr int FANCI_MAXOCC = 30;  // occupied pairs per determinant (2^m Ryser terms)

// (-1)^m sum_{S subset of the m columns} (-1)^|S| prod_i sum_{j in S} a(i, j), in the
// reference's order (ap1rog.cpp:104-130): the permanent of the m x m matrix a.
template <typename A>
__device__ inline double ryser(int m, const A &a) {
    double out = 0.0;
    const unsigned long long c = 1ULL << m;
    for (unsigned long long k = 0; k < c; ++k) {
        double prod = 1.0;
        for (int i = 0; i < m; ++i) {
            double rowsum = 0.0;
            for (int j = 0; j < m; ++j)
                if (k & (1ULL << j)) rowsum += a(i, j);
            prod *= rowsum;
        }
        // __popcll is unsigned: branch on the parity rather than form 1 - 2 * parity
        out += (__popcll(k) & 1u) ? -prod : prod;
    }
    return out * ((m % 2 == 1) ? -1.0 : 1.0);
}

// Per-determinant lists (init_overlap): AP1roG: holes (HF orbitals emptied) and
// particles (virtual index p - nocc), with their count; APIG: the occupied orbitals.
struct FanciView {
    int kind, nocc, nbasis, ncol;  // ncol: row stride of the coefficient matrix
    i64 nconn, nparam;
    const int *holes, *parts, *nexc;  // nconn x nocc, nconn x nocc, nconn
};

__global__ void k_fanci_overlap(FanciView F, i64 ndet, const double *__restrict__ x, double *__restrict__ y) {
    const i64 d = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndet) return;
    const int *h = F.holes + d * F.nocc, *p = F.parts + d * F.nocc;
    if (F.kind == PYCI_FANCI_AP1ROG) {
        const int m = F.nexc[d];
        if (m == 0) {
            y[d] = 1.0;
            return;
        }
        y[d] = ryser(m, [&](int i, int j) { return x[F.ncol * h[i] + p[j]]; });
    } else {
        // APIG (apig.cpp:87-125): rows = occupied orbitals, columns = geminals
        const int m = F.nocc;
        y[d] = ryser(m, [&](int i, int j) { return x[F.nocc * p[i] + j]; });
    }
}

// d(c_d)/d(p_k) into Dr[d * nparam + k] (row-major, the SpMM's layout).
__global__ void k_fanci_d_overlap(FanciView F, i64 ndet, const double *__restrict__ x, double *__restrict__ Dr) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ndet * F.nparam) return;
    const i64 d = t / F.nparam, k = t - d * F.nparam;
    const int *h = F.holes + d * F.nocc, *p = F.parts + d * F.nocc;
    int rows[FANCI_MAXOCC], cols[FANCI_MAXOCC];
    int m = 0, n = 0, mfull;
    if (F.kind == PYCI_FANCI_AP1ROG) {
        // ap1rog.cpp:150-186: drop the parameter's hole row and particle column
        mfull = F.nexc[d];
        if (mfull == 0) {
            Dr[t] = 0.0;
            return;
        }
        const int kr = (int)(k / F.ncol), kc = (int)(k % F.ncol);
        for (int i = 0; i < mfull; ++i) {
            if (h[i] != kr) rows[m++] = h[i];
            if (p[i] != kc) cols[n++] = p[i];
        }
    } else {
        // apig.cpp:137-170: drop geminal k % nocc and orbital k / nocc
        mfull = F.nocc;
        const int kr = (int)(k % F.nocc), kc = (int)(k / F.nocc);
        for (int i = 0; i < mfull; ++i) {
            if (i != kr) rows[m++] = i;
            if (p[i] != kc) cols[n++] = p[i];
        }
    }
    if (m == 0 && n == 0) {
        Dr[t] = 1.0;
        return;
    }
    if (m == mfull || n == mfull || m != n) {
        Dr[t] = 0.0;
        return;
    }
    if (F.kind == PYCI_FANCI_AP1ROG)
        Dr[t] = ryser(m, [&](int i, int j) { return x[F.ncol * rows[i] + cols[j]]; });
    else
        Dr[t] = ryser(m, [&](int i, int j) { return x[F.nocc * cols[i] + rows[j]]; });
}

// f[i] = (H c)[i] - e c[i] (i < nproj), then the determinant constraints
// c[idx] - val and the parameter constraints x[idx] - val (fanci.cpp:129-160).
__global__ void k_fanci_objective(i64 nproj, double e, const double *hc, const double *c, const double *x,
                                  int ndc, const i64 *idc, const double *vdc, int npc, const i64 *ipc,
                                  const double *vpc, double *f) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nproj) f[i] = hc[i] - e * c[i];
    else if (i < nproj + ndc) f[i] = c[idc[i - nproj]] - vdc[i - nproj];
    else if (i < nproj + ndc + npc) f[i] = x[ipc[i - nproj - ndc]] - vpc[i - nproj - ndc];
}

// Jacobian (column-major, neq x (nparam + 1), fanci.cpp:162-217): column k < nparam
// holds (H D)[:, k] - e D[:nproj, k], the determinant-constraint rows D[idx, k] and
// the parameter-constraint rows delta; the last column -c[:nproj] and zeros.
__global__ void k_fanci_jacobian(i64 nproj, i64 nparam, double e, const double *HD, const double *Dr, const double *c,
                                 int ndc, const i64 *idc, int npc, const i64 *ipc, double *J) {
    const i64 neq = nproj + ndc + npc;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= neq * (nparam + 1)) return;
    const i64 k = t / neq, r = t - k * neq;
    double v;
    if (k == nparam) {
        v = (r < nproj) ? -c[r] : 0.0;
    } else if (r < nproj) {
        v = HD[r * nparam + k] - e * Dr[r * nparam + k];
    } else if (r < nproj + ndc) {
        v = Dr[idc[r - nproj] * nparam + k];
    } else {
        v = (ipc[r - nproj - ndc] == k) ? 1.0 : 0.0;
    }
    J[t] = v;
}

// d_overlap output of the reference (column-major nconn x nparam) from Dr.
__global__ void k_row2colmajor(i64 nrow, i64 ncol, const double *src, double *dst) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow * ncol) return;
    const i64 k = t / nrow, r = t - k * nrow;
    dst[t] = src[r * ncol + k];
}

}  // namespace pe

using namespace pe;

struct pyci_fanci {
    pyci_op *op = nullptr;  // not owned: nproj x nconn, symmetric = 0
    int device = 0;         // the operator's device (destroy does not read op)
    FanciView view{};
    i64 nproj = 0, nconn = 0, nparam = 0;
    DevBuf<int> holes, parts, nexc;
    std::vector<i64> idc, ipc;
    std::vector<double> vdc, vpc;
    DevBuf<i64> didc, dipc;
    DevBuf<double> dvdc, dvpc, x, c, hc, f, Dr, HD, J;
};

namespace pe {

pyci_fanci *fanci_create(pyci_op *op, pyci_wfn *wfn, int kind, i64 ndc, const i64 *idc, const double *vdc, i64 npc,
                         const i64 *ipc, const double *vpc) {
    PE_REQUIRE(kind == PYCI_FANCI_AP1ROG || kind == PYCI_FANCI_APIG, PYCI_EINVAL, "unknown FanCI objective");
    PE_REQUIRE(wfn->kind == PYCI_WFN_DOCI, PYCI_ETYPE, "the geminal objectives take a DOCI wave function");
    PE_REQUIRE(op->storage == PYCI_STORAGE_CSR && !op->symmetric, PYCI_EINVAL,
               "the FanCI operator must be a non-symmetric (nproj x nconn) CSR operator");
    PE_REQUIRE(op->ncol <= wfn->ndet && op->nrow <= op->ncol, PYCI_EINVAL, "operator and wave function do not match");
    const int nocc = (int)wfn->nup, nb = (int)wfn->nbasis;
    PE_REQUIRE(nocc >= 1 && nocc <= FANCI_MAXOCC, PYCI_EUNSUPPORTED, "FanCI overlaps take 1..30 occupied pairs");
    auto F = std::make_unique<pyci_fanci>();
    F->op = op;
    F->device = op->ctx->device;
    F->nproj = op->nrow;
    F->nconn = op->ncol;
    F->nparam = (kind == PYCI_FANCI_AP1ROG) ? (i64)nocc * (nb - nocc) : (i64)nocc * nb;
    const i64 nconn = F->nconn;
    // hole / particle lists of the first nconn determinants (init_overlap)
    std::vector<u64> dets((size_t)nconn);
    PE_HIP(hipMemcpy(dets.data(), wfn->dets.p, sizeof(u64) * nconn, hipMemcpyDeviceToHost));
    std::vector<int> hl((size_t)(nconn * nocc), 0), pl((size_t)(nconn * nocc), 0), ne((size_t)nconn, 0);
    const u64 hf = (nocc >= 64) ? ~0ULL : ((1ULL << nocc) - 1ULL);
    for (i64 d = 0; d < nconn; ++d) {
        const u64 det = dets[(size_t)d];
        if (kind == PYCI_FANCI_AP1ROG) {
            u64 hw = (hf ^ det) & hf, pw = (hf ^ det) & det;
            int e = 0;
            while (hw) {
                hl[(size_t)(d * nocc + e)] = __builtin_ctzll(hw);
                pl[(size_t)(d * nocc + e)] = __builtin_ctzll(pw) - nocc;
                hw &= hw - 1;
                pw &= pw - 1;
                ++e;
            }
            ne[(size_t)d] = e;
        } else {
            int e = 0;
            for (u64 w = det; w; w &= w - 1) pl[(size_t)(d * nocc + e++)] = __builtin_ctzll(w);
            ne[(size_t)d] = e;
        }
    }
    F->holes.alloc(hl.size());
    F->parts.alloc(pl.size());
    F->nexc.alloc(ne.size());
    PE_HIP(hipMemcpy(F->holes.p, hl.data(), sizeof(int) * hl.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(F->parts.p, pl.data(), sizeof(int) * pl.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(F->nexc.p, ne.data(), sizeof(int) * ne.size(), hipMemcpyHostToDevice));
    F->view = FanciView{kind, nocc, nb, nb - nocc, nconn, F->nparam, F->holes.p, F->parts.p, F->nexc.p};
    // constraints
    for (i64 i = 0; i < ndc; ++i) {
        PE_REQUIRE(idc[i] >= 0 && idc[i] < nconn, PYCI_EINVAL, "determinant constraint index out of range");
        F->idc.push_back(idc[i]);
        F->vdc.push_back(vdc[i]);
    }
    for (i64 i = 0; i < npc; ++i) {
        PE_REQUIRE(ipc[i] >= 0 && ipc[i] <= F->nparam, PYCI_EINVAL, "parameter constraint index out of range");
        F->ipc.push_back(ipc[i]);
        F->vpc.push_back(vpc[i]);
    }
    F->didc.alloc(std::max<size_t>(1, F->idc.size()));
    F->dipc.alloc(std::max<size_t>(1, F->ipc.size()));
    F->dvdc.alloc(std::max<size_t>(1, F->vdc.size()));
    F->dvpc.alloc(std::max<size_t>(1, F->vpc.size()));
    if (ndc) {
        PE_HIP(hipMemcpy(F->didc.p, F->idc.data(), sizeof(i64) * ndc, hipMemcpyHostToDevice));
        PE_HIP(hipMemcpy(F->dvdc.p, F->vdc.data(), sizeof(double) * ndc, hipMemcpyHostToDevice));
    }
    if (npc) {
        PE_HIP(hipMemcpy(F->dipc.p, F->ipc.data(), sizeof(i64) * npc, hipMemcpyHostToDevice));
        PE_HIP(hipMemcpy(F->dvpc.p, F->vpc.data(), sizeof(double) * npc, hipMemcpyHostToDevice));
    }
    const i64 neq = F->nproj + ndc + npc;
    F->x.alloc((size_t)(F->nparam + 1));
    F->c.alloc((size_t)nconn);
    F->hc.alloc((size_t)F->nproj);
    F->f.alloc((size_t)neq);
    F->Dr.alloc((size_t)(nconn * std::max<i64>(F->nparam, 1)));
    F->HD.alloc((size_t)(F->nproj * std::max<i64>(F->nparam, 1)));
    // J also receives the column-major d_overlap output (nconn x nparam)
    F->J.alloc((size_t)std::max<i64>(neq * (F->nparam + 1), nconn * F->nparam));
    return F.release();
}

int fanci_device(const pyci_fanci *F) { return F->device; }

void fanci_destroy(pyci_fanci *F) { delete F; }

void fanci_info(const pyci_fanci *F, i64 *nparam, i64 *nproj, i64 *nconn, i64 *nequation) {
    if (nparam) *nparam = F->nparam;
    if (nproj) *nproj = F->nproj;
    if (nconn) *nconn = F->nconn;
    if (nequation) *nequation = F->nproj + (i64)F->idc.size() + (i64)F->ipc.size();
}

static void upload_params(pyci_fanci *F, const double *x, i64 n, hipStream_t s) {
    PE_HIP(hipMemcpyAsync(F->x.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
}

void fanci_overlap(pyci_fanci *F, const double *x, double *y) {
    hipStream_t s = F->op->ctx->stream;
    upload_params(F, x, F->nparam, s);
    k_fanci_overlap<<<grid_for(F->nconn, 256), 256, 0, s>>>(F->view, F->nconn, F->x.p, F->c.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(y, F->c.p, sizeof(double) * F->nconn, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

void fanci_d_overlap(pyci_fanci *F, const double *x, double *y) {
    hipStream_t s = F->op->ctx->stream;
    upload_params(F, x, F->nparam, s);
    const i64 n = F->nconn * F->nparam;
    if (n > 0) {
        k_fanci_d_overlap<<<grid_for(n, 256), 256, 0, s>>>(F->view, F->nconn, F->x.p, F->Dr.p);
        PE_LAUNCH_CHECK();
        k_row2colmajor<<<grid_for(n, 256), 256, 0, s>>>(F->nconn, F->nparam, F->Dr.p, F->J.p);
        PE_LAUNCH_CHECK();
        PE_HIP(hipMemcpyAsync(y, F->J.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    PE_HIP(hipStreamSynchronize(s));
}

void fanci_objective(pyci_fanci *F, const double *x, double *y) {
    hipStream_t s = F->op->ctx->stream;
    upload_params(F, x, F->nparam + 1, s);
    k_fanci_overlap<<<grid_for(F->nconn, 256), 256, 0, s>>>(F->view, F->nconn, F->x.p, F->c.p);
    PE_LAUNCH_CHECK();
    sigma_f64(F->op, 1, F->c.p, F->hc.p, s);
    const int ndc = (int)F->idc.size(), npc = (int)F->ipc.size();
    const i64 neq = F->nproj + ndc + npc;
    k_fanci_objective<<<grid_for(neq, 256), 256, 0, s>>>(F->nproj, x[F->nparam], F->hc.p, F->c.p, F->x.p, ndc,
                                                         F->didc.p, F->dvdc.p, npc, F->dipc.p, F->dvpc.p, F->f.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(y, F->f.p, sizeof(double) * neq, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

void fanci_jacobian(pyci_fanci *F, const double *x, double *y) {
    hipStream_t s = F->op->ctx->stream;
    upload_params(F, x, F->nparam + 1, s);
    // overlaps of the projection space only (fanci.cpp:169)
    k_fanci_overlap<<<grid_for(F->nproj, 256), 256, 0, s>>>(F->view, F->nproj, F->x.p, F->c.p);
    PE_LAUNCH_CHECK();
    const i64 n = F->nconn * F->nparam;
    if (n > 0) {
        k_fanci_d_overlap<<<grid_for(n, 256), 256, 0, s>>>(F->view, F->nconn, F->x.p, F->Dr.p);
        PE_LAUNCH_CHECK();
        // every parameter column through the rectangular operator in one product
        sigma_f64(F->op, (int)F->nparam, F->Dr.p, F->HD.p, s);
    }
    const int ndc = (int)F->idc.size(), npc = (int)F->ipc.size();
    const i64 neq = F->nproj + ndc + npc, tot = neq * (F->nparam + 1);
    k_fanci_jacobian<<<grid_for(tot, 256), 256, 0, s>>>(F->nproj, F->nparam, x[F->nparam], F->HD.p, F->Dr.p, F->c.p,
                                                        ndc, F->didc.p, npc, F->dipc.p, F->J.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(y, F->J.p, sizeof(double) * tot, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

}  // namespace pe
