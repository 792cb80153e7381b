// One- and two-particle reduced density matrices on the GPU.
//
// Reference: compute_rdms for DOCIWfn (rdm.cpp:20-65) and FullCIWfn
// (rdm.cpp:269-530): a loop over determinants I that visits every excitation
// J of I present in the wave function with J > I and scatters c_I c_J times the
// phase into the symmetric index permutations, plus the diagonal c_I^2 terms.
// GenCIWfn is computed from the operator algebra (every ordered pair I, J); the
// reference's GenCI routine (rdm.cpp:532-632) has index errors and is not
// reproduced.
//
// One wavefront per determinant; lanes take excitations from the same flat
// candidate index as the CSR builder. Accumulation uses hardware FP64 atomics
// into R replicated copies of the RDM arrays (block b uses copy b % R) to spread
// contention on hot elements, followed by a reduction over the copies.
#include <algorithm>

#include "internal.hpp"

namespace pe {

__device__ inline void aadd(double *p, double v) { atomicAdd(p, v); }

__device__ inline int nth_bit_r(u64 x, int t) {
    for (int s = 0; s < t && x; ++s) x &= x - 1;
    return x ? __builtin_ctzll(x) : 0;
}

__device__ inline void tri_decode_r(int t, int &i, int &k) {
    int kk = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
    while (kk * (kk - 1) / 2 > t) --kk;
    while ((kk + 1) * kk / 2 <= t) ++kk;
    k = kk;
    i = t - kk * (kk - 1) / 2;
}

// 8 antisymmetric images of a same-spin 2-body term X[p,q,r,s] for pair (p,q),(r,s) and its transpose
__device__ inline void same_spin8(double *X, int n, int ii, int kk, int jj, int ll, double v) {
    const i64 n1 = n, n2 = n1 * n1, n3 = n2 * n1;
    aadd(&X[ii * n3 + kk * n2 + jj * n1 + ll], v);
    aadd(&X[ii * n3 + kk * n2 + ll * n1 + jj], -v);
    aadd(&X[kk * n3 + ii * n2 + jj * n1 + ll], -v);
    aadd(&X[kk * n3 + ii * n2 + ll * n1 + jj], v);
    aadd(&X[jj * n3 + ll * n2 + ii * n1 + kk], v);
    aadd(&X[jj * n3 + ll * n2 + kk * n1 + ii], -v);
    aadd(&X[ll * n3 + jj * n2 + ii * n1 + kk], -v);
    aadd(&X[ll * n3 + jj * n2 + kk * n1 + ii], v);
}

// single i->j with spectator k of the same spin (rdm.cpp:331-349 pattern)
__device__ inline void same_spin_single(double *X, int n, int ii, int jj, int kk, double v) {
    const i64 n1 = n, n2 = n1 * n1, n3 = n2 * n1;
    aadd(&X[ii * n3 + kk * n2 + jj * n1 + kk], v);
    aadd(&X[ii * n3 + kk * n2 + kk * n1 + jj], -v);
    aadd(&X[kk * n3 + ii * n2 + kk * n1 + jj], v);
    aadd(&X[kk * n3 + ii * n2 + jj * n1 + kk], -v);
    aadd(&X[jj * n3 + kk * n2 + ii * n1 + kk], v);
    aadd(&X[jj * n3 + kk * n2 + kk * n1 + ii], -v);
    aadd(&X[kk * n3 + jj * n2 + ii * n1 + kk], -v);
    aadd(&X[kk * n3 + jj * n2 + kk * n1 + ii], v);
}

// Flat candidate index of one row (the CSR builder's order): alpha singles, 1-1
// doubles, alpha doubles, beta singles, beta doubles; c0..c4 are the class ends.
template <int KIND>
__device__ inline void exc_bounds(int nou, int nvu, int nod, int nvd, int &c0, int &c1, int &c2, int &c3, int &c4) {
    const int nsu = nou * nvu, nsd = nod * nvd;
    const int npv_u = nvu * (nvu - 1) / 2, npv_d = nvd * (nvd - 1) / 2;
    if (KIND == PYCI_WFN_FULLCI) {
        c0 = nsu;
        c1 = c0 + nsu * nsd;
        c2 = c1 + (nou * (nou - 1) / 2) * npv_u;
        c3 = c2 + nsd;
        c4 = c3 + (nod * (nod - 1) / 2) * npv_d;
    } else if (KIND == PYCI_WFN_DOCI) {
        c0 = c1 = c2 = c3 = c4 = nsu;
    } else {
        c0 = c1 = nsu;
        c2 = c3 = c4 = c1 + (nou * (nou - 1) / 2) * npv_u;
    }
}

// Candidate e of a row: its class and list positions (i, a: first occupied and
// virtual; k, l: second), in the lists each lane holds in `packed`.
__device__ inline void exc_decode(int e, int c0, int c1, int c2, int c3, int nvu, int nod, int nvd, int &cls, int &i,
                                  int &a, int &k, int &l) {
    const int nsd = nod * nvd, npv_u = nvu * (nvu - 1) / 2, npv_d = nvd * (nvd - 1) / 2;
    cls = i = a = k = l = 0;
    if (e < c0) {
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
        tri_decode_r(po, i, k);
        tri_decode_r(pv, a, l);
    } else if (e < c3) {
        cls = 3;
        int t = e - c2;
        i = t / nvd;
        a = t - i * nvd;
    } else {
        cls = 4;
        int t = e - c3, po = t / npv_d, pv = t - po * npv_d;
        tri_decode_r(po, i, k);
        tri_decode_r(pv, a, l);
    }
}

template <int KIND>
__global__ void __launch_bounds__(256) k_rdm(WfnView w, const double *c, int R, i64 stride, double *acc) {
    const int lane = threadIdx.x & 63;
    const i64 row = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= w.ndet) return;
    const int n = w.nb;
    const i64 n1 = n, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    double *base = acc + (i64)(blockIdx.x % R) * stride;
    const u64 m = full_mask(n);
    const u64 up = w.dets[(i64)w.nkey * row];
    const u64 dn = (KIND == PYCI_WFN_FULLCI) ? w.dets[2 * row + 1] : 0ULL;
    const int nou = __popcll(up), nvu = n - nou, nod = __popcll(dn), nvd = n - nod;
    const int packed = nth_bit_r(up, lane) | (nth_bit_r(~up & m, lane) << 8) | (nth_bit_r(dn, lane) << 16) |
                       (nth_bit_r(~dn & m, lane) << 24);
    const double ci = c[row];
    // ---- diagonal terms (lane-parallel over occupied pairs) ----
    const double val1 = ci * ci;
    if (KIND == PYCI_WFN_DOCI) {
        double *d0 = base, *d2 = base + n2;
        for (int t = lane; t < nou * nou; t += 64) {
            const int a = t / nou, b = t - a * nou;
            const int p = nth_bit_r(up, a), q = nth_bit_r(up, b);
            if (a == b)
                aadd(&d0[p * (n + 1)], val1);
            else
                aadd(&d2[p * n1 + q], val1);
        }
    } else if (KIND == PYCI_WFN_FULLCI) {
        double *aa = base, *bb = base + n2, *aaaa = base + 2 * n2, *bbbb = aaaa + n4, *abab = bbbb + n4;
        for (int t = lane; t < nou * nou; t += 64) {
            const int a = t / nou, b = t - a * nou;
            const int p = nth_bit_r(up, a), q = nth_bit_r(up, b);
            if (a == b) {
                aadd(&aa[p * (n + 1)], val1);
            } else if (a < b) {
                aadd(&aaaa[p * n3 + q * n2 + p * n1 + q], val1);
                aadd(&aaaa[p * n3 + q * n2 + q * n1 + p], -val1);
                aadd(&aaaa[q * n3 + p * n2 + p * n1 + q], -val1);
                aadd(&aaaa[q * n3 + p * n2 + q * n1 + p], val1);
            }
        }
        for (int t = lane; t < nod * nod; t += 64) {
            const int a = t / nod, b = t - a * nod;
            const int p = nth_bit_r(dn, a), q = nth_bit_r(dn, b);
            if (a == b) {
                aadd(&bb[p * (n + 1)], val1);
            } else if (a < b) {
                aadd(&bbbb[p * n3 + q * n2 + p * n1 + q], val1);
                aadd(&bbbb[p * n3 + q * n2 + q * n1 + p], -val1);
                aadd(&bbbb[q * n3 + p * n2 + p * n1 + q], -val1);
                aadd(&bbbb[q * n3 + p * n2 + q * n1 + p], val1);
            }
        }
        for (int t = lane; t < nou * nod; t += 64) {
            const int a = t / nod, b = t - a * nod;
            const int p = nth_bit_r(up, a), q = nth_bit_r(dn, b);
            aadd(&abab[p * n3 + q * n2 + p * n1 + q], val1);
        }
    } else {
        double *r1 = base, *r2 = base + n2;
        for (int t = lane; t < nou * nou; t += 64) {
            const int a = t / nou, b = t - a * nou;
            const int p = nth_bit_r(up, a), q = nth_bit_r(up, b);
            if (a == b) {
                aadd(&r1[p * (n + 1)], val1);
            } else {
                aadd(&r2[p * n3 + q * n2 + p * n1 + q], val1);
                aadd(&r2[p * n3 + q * n2 + q * n1 + p], -val1);
            }
        }
    }
    // ---- off-diagonal terms ----
    int c0, c1, c2, c3, c4;
    exc_bounds<KIND>(nou, nvu, nod, nvd, c0, c1, c2, c3, c4);
    for (int e0 = 0; e0 < c4; e0 += 64) {
        const int e = e0 + lane;
        const bool valid = e < c4;
        int cls = 0, i = 0, a = 0, k = 0, l = 0;
        if (valid) exc_decode(e, c0, c1, c2, c3, nvu, nod, nvd, cls, i, a, k, l);
        const int pi = __shfl(packed, i), pa = __shfl(packed, a), pk = __shfl(packed, k), pl = __shfl(packed, l);
        if (!valid) continue;
        const bool alpha = cls <= 2;
        const int ii = alpha ? (pi & 255) : ((pi >> 16) & 255);
        const int jj = alpha ? ((pa >> 8) & 255) : ((pa >> 24) & 255);
        int kk, ll;
        if (cls == 1 || !alpha) {
            kk = (pk >> 16) & 255;
            ll = (pl >> 24) & 255;
        } else {
            kk = pk & 255;
            ll = (pl >> 8) & 255;
        }
        u64 nu = up, nd = dn;
        if (cls == 0) nu = up ^ (1ULL << ii) ^ (1ULL << jj);
        else if (cls == 1) {
            nu = up ^ (1ULL << ii) ^ (1ULL << jj);
            nd = dn ^ (1ULL << kk) ^ (1ULL << ll);
        } else if (cls == 2) nu = up ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
        else if (cls == 3) nd = dn ^ (1ULL << ii) ^ (1ULL << jj);
        else nd = dn ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
        const i64 col = wfn_lookup(w, nu, nd);
        if (col < 0 || col >= w.ndet) continue;
        if (KIND != PYCI_WFN_GENCI && col <= row) continue;
        const double cc = ci * c[col];
        if (KIND == PYCI_WFN_DOCI) {
            double *d0 = base;
            aadd(&d0[ii * n1 + jj], cc);
            aadd(&d0[jj * n1 + ii], cc);
        } else if (KIND == PYCI_WFN_FULLCI) {
            double *aa = base, *bb = base + n2, *aaaa = base + 2 * n2, *bbbb = aaaa + n4, *abab = bbbb + n4;
            if (cls == 0) {
                const double v = cc * phase1(up, ii, jj);
                aadd(&aa[ii * n1 + jj], v);
                aadd(&aa[jj * n1 + ii], v);
                for (u64 x = up; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    if (q != ii) same_spin_single(aaaa, n, ii, jj, q, v);
                }
                for (u64 x = dn; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    aadd(&abab[ii * n3 + q * n2 + jj * n1 + q], v);
                    aadd(&abab[jj * n3 + q * n2 + ii * n1 + q], v);
                }
            } else if (cls == 1) {
                const double v = cc * (double)(phase1(up, ii, jj) * phase1(dn, kk, ll));
                aadd(&abab[ii * n3 + kk * n2 + jj * n1 + ll], v);
                aadd(&abab[jj * n3 + ll * n2 + ii * n1 + kk], v);
            } else if (cls == 2) {
                same_spin8(aaaa, n, ii, kk, jj, ll, cc * phase2(up, ii, kk, jj, ll));
            } else if (cls == 3) {
                const double v = cc * phase1(dn, ii, jj);
                aadd(&bb[ii * n1 + jj], v);
                aadd(&bb[jj * n1 + ii], v);
                for (u64 x = up; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    aadd(&abab[q * n3 + ii * n2 + q * n1 + jj], v);
                    aadd(&abab[q * n3 + jj * n2 + q * n1 + ii], v);
                }
                for (u64 x = dn; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    if (q != ii) same_spin_single(bbbb, n, ii, jj, q, v);
                }
            } else {
                same_spin8(bbbb, n, ii, kk, jj, ll, cc * phase2(dn, ii, kk, jj, ll));
            }
        } else {
            double *r1 = base, *r2 = base + n2;
            if (cls == 0) {
                // ket J = this row, bra I = col: <I| a+_jj a_ii |J>
                const double v = cc * phase1(up, ii, jj);
                aadd(&r1[jj * n1 + ii], v);
                for (u64 x = up; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    if (q == ii) continue;
                    aadd(&r2[jj * n3 + q * n2 + ii * n1 + q], v);
                    aadd(&r2[q * n3 + jj * n2 + q * n1 + ii], v);
                    aadd(&r2[jj * n3 + q * n2 + q * n1 + ii], -v);
                    aadd(&r2[q * n3 + jj * n2 + ii * n1 + q], -v);
                }
            } else {
                const double v = cc * phase2(up, ii, kk, jj, ll);
                aadd(&r2[jj * n3 + ll * n2 + ii * n1 + kk], v);
                aadd(&r2[jj * n3 + ll * n2 + kk * n1 + ii], -v);
                aadd(&r2[ll * n3 + jj * n2 + ii * n1 + kk], -v);
                aadd(&r2[ll * n3 + jj * n2 + kk * n1 + ii], v);
            }
        }
    }
}

// Transition RDMs (rdm.cpp:634-899): every determinant I of set 1 (w1, c1) and
// its excitations J present in set 2 (w2, c2) add c1_I c2_J times the phase to
// the reference's index images, one term per ordered (I, J) -- no symmetric
// images, as <1|...|2> is not symmetric. Diagonal: I itself where it is in set 2.
// Element conventions: [p,q] = <1| a+_p a_q |2>, [p,q,r,s] = <1| a+_p a+_q a_s a_r |2>;
// GenCI by the operator algebra (the reference routine has index errors).
template <int KIND>
__global__ void __launch_bounds__(256) k_trdm(WfnView w1, WfnView w2, const double *c1, const double *c2, int R,
                                             i64 stride, double *acc) {
    const int lane = threadIdx.x & 63;
    const i64 row = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= w1.ndet) return;
    const int n = w1.nb;
    const i64 n1 = n, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    double *base = acc + (i64)(blockIdx.x % R) * stride;
    const u64 m = full_mask(n);
    const u64 up = w1.dets[(i64)w1.nkey * row];
    const u64 dn = (KIND == PYCI_WFN_FULLCI) ? w1.dets[2 * row + 1] : 0ULL;
    const int nou = __popcll(up), nvu = n - nou, nod = __popcll(dn), nvd = n - nod;
    const int packed = nth_bit_r(up, lane) | (nth_bit_r(~up & m, lane) << 8) | (nth_bit_r(dn, lane) << 16) |
                       (nth_bit_r(~dn & m, lane) << 24);
    const double ci = c1[row];
    // the 4 antisymmetric images of a same-spin pair term (p,q | r,s)
    auto four = [&](double *X, int p, int q, int r, int t, double v) {
        aadd(&X[p * n3 + q * n2 + r * n1 + t], v);
        aadd(&X[p * n3 + q * n2 + t * n1 + r], -v);
        aadd(&X[q * n3 + p * n2 + r * n1 + t], -v);
        aadd(&X[q * n3 + p * n2 + t * n1 + r], v);
    };
    // ---- diagonal terms ----
    const i64 jd = wfn_lookup(w2, up, dn);
    if (jd >= 0 && jd < w2.ndet) {
        const double val1 = ci * c2[jd];
        if (KIND == PYCI_WFN_DOCI) {
            double *d0 = base, *d2 = base + n2;
            for (int t = lane; t < nou * nou; t += 64) {
                const int a = t / nou, b = t - a * nou;
                const int p = nth_bit_r(up, a), q = nth_bit_r(up, b);
                if (a == b)
                    aadd(&d0[p * (n + 1)], val1);
                else
                    aadd(&d2[p * n1 + q], val1);
            }
        } else {
            double *r1a = base, *r2a = base + (KIND == PYCI_WFN_FULLCI ? 2 * n2 : n2);
            for (int t = lane; t < nou * nou; t += 64) {
                const int a = t / nou, b = t - a * nou;
                const int p = nth_bit_r(up, a), q = nth_bit_r(up, b);
                if (a == b)
                    aadd(&r1a[p * (n + 1)], val1);
                else if (a < b)
                    four(r2a, p, q, p, q, val1);
            }
            if (KIND == PYCI_WFN_FULLCI) {
                double *bb = base + n2, *bbbb = base + 2 * n2 + n4, *abab = bbbb + n4;
                for (int t = lane; t < nod * nod; t += 64) {
                    const int a = t / nod, b = t - a * nod;
                    const int p = nth_bit_r(dn, a), q = nth_bit_r(dn, b);
                    if (a == b)
                        aadd(&bb[p * (n + 1)], val1);
                    else if (a < b)
                        four(bbbb, p, q, p, q, val1);
                }
                for (int t = lane; t < nou * nod; t += 64) {
                    const int a = t / nod, b = t - a * nod;
                    const int p = nth_bit_r(up, a), q = nth_bit_r(dn, b);
                    aadd(&abab[p * n3 + q * n2 + p * n1 + q], val1);
                }
            }
        }
    }
    // ---- excitations of I present in set 2 ----
    int c0, cb1, cb2, cb3, c4;
    exc_bounds<KIND>(nou, nvu, nod, nvd, c0, cb1, cb2, cb3, c4);
    for (int e0 = 0; e0 < c4; e0 += 64) {
        const int e = e0 + lane;
        const bool valid = e < c4;
        int cls = 0, i = 0, a = 0, k = 0, l = 0;
        if (valid) exc_decode(e, c0, cb1, cb2, cb3, nvu, nod, nvd, cls, i, a, k, l);
        const int pi = __shfl(packed, i), pa = __shfl(packed, a), pk = __shfl(packed, k), pl = __shfl(packed, l);
        if (!valid) continue;
        const bool alpha = cls <= 2;
        const int ii = alpha ? (pi & 255) : ((pi >> 16) & 255);
        const int jj = alpha ? ((pa >> 8) & 255) : ((pa >> 24) & 255);
        int kk, ll;
        if (cls == 1 || !alpha) {
            kk = (pk >> 16) & 255;
            ll = (pl >> 24) & 255;
        } else {
            kk = pk & 255;
            ll = (pl >> 8) & 255;
        }
        u64 nu = up, nd = dn;
        if (KIND == PYCI_WFN_DOCI) nu = up ^ (1ULL << ii) ^ (1ULL << jj);
        else if (cls == 0) nu = up ^ (1ULL << ii) ^ (1ULL << jj);
        else if (cls == 1) {
            nu = up ^ (1ULL << ii) ^ (1ULL << jj);
            nd = dn ^ (1ULL << kk) ^ (1ULL << ll);
        } else if (cls == 2) nu = up ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
        else if (cls == 3) nd = dn ^ (1ULL << ii) ^ (1ULL << jj);
        else nd = dn ^ (1ULL << ii) ^ (1ULL << jj) ^ (1ULL << kk) ^ (1ULL << ll);
        const i64 col = wfn_lookup(w2, nu, nd);
        if (col < 0 || col >= w2.ndet) continue;
        const double cc = ci * c2[col];
        if (KIND == PYCI_WFN_DOCI) {
            aadd(&base[ii * n1 + jj], cc);
        } else if (KIND == PYCI_WFN_GENCI || cls <= 2) {
            // alpha string (FullCI) or the only string (GenCI)
            double *r1 = base, *r2 = base + (KIND == PYCI_WFN_FULLCI ? 2 * n2 : n2);
            double *abab = base + 2 * n2 + 2 * n4;
            if (cls == 0) {
                const double v = cc * phase1(up, ii, jj);
                aadd(&r1[ii * n1 + jj], v);
                for (u64 x = up; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    if (q == ii) continue;
                    aadd(&r2[ii * n3 + q * n2 + jj * n1 + q], v);
                    aadd(&r2[ii * n3 + q * n2 + q * n1 + jj], -v);
                    aadd(&r2[q * n3 + ii * n2 + q * n1 + jj], v);
                    aadd(&r2[q * n3 + ii * n2 + jj * n1 + q], -v);
                }
                if (KIND == PYCI_WFN_FULLCI)
                    for (u64 x = dn; x; x &= x - 1) {
                        const int q = __builtin_ctzll(x);
                        aadd(&abab[ii * n3 + q * n2 + jj * n1 + q], v);
                    }
            } else if (cls == 1) {
                aadd(&abab[ii * n3 + kk * n2 + jj * n1 + ll], cc * (double)(phase1(up, ii, jj) * phase1(dn, kk, ll)));
            } else {
                four(r2, ii, kk, jj, ll, cc * phase2(up, ii, kk, jj, ll));
            }
        } else {
            double *bb = base + n2, *bbbb = base + 2 * n2 + n4, *abab = bbbb + n4;
            if (cls == 3) {
                const double v = cc * phase1(dn, ii, jj);
                aadd(&bb[ii * n1 + jj], v);
                for (u64 x = up; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    aadd(&abab[q * n3 + ii * n2 + q * n1 + jj], v);
                }
                for (u64 x = dn; x; x &= x - 1) {
                    const int q = __builtin_ctzll(x);
                    if (q == ii) continue;
                    aadd(&bbbb[ii * n3 + q * n2 + jj * n1 + q], v);
                    aadd(&bbbb[ii * n3 + q * n2 + q * n1 + jj], -v);
                    aadd(&bbbb[q * n3 + ii * n2 + q * n1 + jj], v);
                    aadd(&bbbb[q * n3 + ii * n2 + jj * n1 + q], -v);
                }
            } else {
                four(bbbb, ii, kk, jj, ll, cc * phase2(dn, ii, kk, jj, ll));
            }
        }
    }
}

__global__ void k_sum_copies(const double *acc, int R, i64 stride, double *out) {
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < stride; t += (i64)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int r = 0; r < R; ++r) v += acc[(i64)r * stride + t];
        out[t] = v;
    }
}

// Sizes of the two output arrays of a kind (rdm1, rdm2 element counts).
static void rdm_sizes(int kind, i64 n, i64 &s1, i64 &s2) {
    const i64 n2 = n * n, n4 = n2 * n2;
    if (kind == PYCI_WFN_DOCI) {
        s1 = n2;
        s2 = n2;
    } else if (kind == PYCI_WFN_FULLCI) {
        s1 = 2 * n2;
        s2 = 3 * n4;
    } else {
        s1 = n2;
        s2 = n4;
    }
}

void trdms_device(pyci_wfn *w1, pyci_wfn *w2, const double *c1_host, const double *c2_host, double *rdm1,
                  double *rdm2) {
    PE_REQUIRE(w1->ctx == w2->ctx, PYCI_EINVAL, "wave functions belong to different contexts");
    PE_REQUIRE(w1->kind == w2->kind && w1->nbasis == w2->nbasis && w1->nup == w2->nup && w1->ndn == w2->ndn,
               PYCI_EINVAL, "transition RDMs need two wave functions of the same kind, nbasis and occupations");
    hipStream_t s = w1->ctx->stream;
    i64 s1, s2;
    rdm_sizes(w1->kind, w1->nbasis, s1, s2);
    const i64 stride = s1 + s2;
    const i64 nd1 = w1->ndet, nd2 = w2->ndet;
    int R = (int)std::max<i64>(1, std::min<i64>(32, ((i64)256 << 20) / (stride * 8)));
    R = (int)std::min<i64>(R, std::max<i64>(1, (nd1 + 3) / 4));
    DevBuf<double> acc((size_t)(R * stride)), out((size_t)stride);
    DevBuf<double> d1((size_t)std::max<i64>(nd1, 1)), d2((size_t)std::max<i64>(nd2, 1));
    PE_HIP(hipMemsetAsync(acc.p, 0, sizeof(double) * R * stride, s));
    if (nd1 > 0 && nd2 > 0) {
        PE_HIP(hipMemcpyAsync(d1.p, c1_host, sizeof(double) * nd1, hipMemcpyHostToDevice, s));
        PE_HIP(hipMemcpyAsync(d2.p, c2_host, sizeof(double) * nd2, hipMemcpyHostToDevice, s));
        const WfnView v1 = view_of(w1), v2 = view_of(w2);
        const unsigned g = grid_for(nd1, 4);
        if (w1->kind == PYCI_WFN_DOCI)
            k_trdm<PYCI_WFN_DOCI><<<g, 256, 0, s>>>(v1, v2, d1.p, d2.p, R, stride, acc.p);
        else if (w1->kind == PYCI_WFN_FULLCI)
            k_trdm<PYCI_WFN_FULLCI><<<g, 256, 0, s>>>(v1, v2, d1.p, d2.p, R, stride, acc.p);
        else
            k_trdm<PYCI_WFN_GENCI><<<g, 256, 0, s>>>(v1, v2, d1.p, d2.p, R, stride, acc.p);
        PE_LAUNCH_CHECK();
    }
    k_sum_copies<<<(unsigned)std::min<i64>(4096, (stride + 255) / 256), 256, 0, s>>>(acc.p, R, stride, out.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(rdm1, out.p, sizeof(double) * s1, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(rdm2, out.p + s1, sizeof(double) * s2, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

void rdms_device(pyci_wfn *wfn, const double *coeffs_host, double *rdm1, double *rdm2) {
    hipStream_t s = wfn->ctx->stream;
    const i64 n = wfn->nbasis, n2 = n * n, n4 = n2 * n2;
    i64 s1, s2;
    if (wfn->kind == PYCI_WFN_DOCI) {
        s1 = n2;
        s2 = n2;
    } else if (wfn->kind == PYCI_WFN_FULLCI) {
        s1 = 2 * n2;
        s2 = 3 * n4;
    } else {
        s1 = n2;
        s2 = n4;
    }
    const i64 stride = s1 + s2;
    const i64 nd = wfn->ndet;
    int R = (int)std::max<i64>(1, std::min<i64>(32, ((i64)256 << 20) / (stride * 8)));
    R = (int)std::min<i64>(R, std::max<i64>(1, (nd + 3) / 4));
    DevBuf<double> acc((size_t)(R * stride)), out((size_t)stride), dc((size_t)std::max<i64>(nd, 1));
    PE_HIP(hipMemsetAsync(acc.p, 0, sizeof(double) * R * stride, s));
    if (nd > 0) {
        PE_HIP(hipMemcpyAsync(dc.p, coeffs_host, sizeof(double) * nd, hipMemcpyHostToDevice, s));
        WfnView w = view_of(wfn);
        unsigned g = grid_for(nd, 4);
        if (wfn->kind == PYCI_WFN_DOCI)
            k_rdm<PYCI_WFN_DOCI><<<g, 256, 0, s>>>(w, dc.p, R, stride, acc.p);
        else if (wfn->kind == PYCI_WFN_FULLCI)
            k_rdm<PYCI_WFN_FULLCI><<<g, 256, 0, s>>>(w, dc.p, R, stride, acc.p);
        else
            k_rdm<PYCI_WFN_GENCI><<<g, 256, 0, s>>>(w, dc.p, R, stride, acc.p);
        PE_LAUNCH_CHECK();
    }
    k_sum_copies<<<(unsigned)std::min<i64>(4096, (stride + 255) / 256), 256, 0, s>>>(acc.p, R, stride, out.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipMemcpyAsync(rdm1, out.p, sizeof(double) * s1, hipMemcpyDeviceToHost, s));
    PE_HIP(hipMemcpyAsync(rdm2, out.p + s1, sizeof(double) * s2, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
}

}  // namespace pe
