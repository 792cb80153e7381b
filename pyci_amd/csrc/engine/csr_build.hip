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
// for the lower-triangle export) and keeps each chunk's hit ballot, an exclusive
// scan turns counts into row pointers, pass 2 writes entries (wave-level ballot
// compaction, probing only the hits) and the diagonal, and each row is sorted by
// column (bitonic in LDS up to 4,096 entries, a segmented radix sort beyond).
// The full FullCI space in colex order skips the count pass and the sort: its
// rows are written in column order from per-string tables (SortedRows).
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace pe {

constexpr int ROWS_PER_BLOCK = 4;  // 4 wavefronts x 64 lanes
constexpr int GENCI_MAXPV = 64 * 63 / 2;  // virtual orbital pairs of a 64-bit GenCI determinant

__device__ inline int nth_bit(u64 x, int t) {
    for (int s = 0; s < t && x; ++s) x &= x - 1;
    return x ? __builtin_ctzll(x) : 0;
}

#define GIX(p, q, r, s) ((((p) * n + (q)) * n + (r)) * n + (s))

// Sorted fill of the full FullCI space in colex order. There every single and
// double excitation of a row is in the space and column = rank(alpha) NB +
// rank(beta), so a row sorted by column is a run of segments, one per string a'
// related to the row's alpha string a, in increasing rank of a': (a, b') for b' in
// the sorted list of b, its singles and doubles (a' = a); (a', b') for b' in the
// sorted list of b and its singles (a' a single of a); (a', b) alone (a' a double of
// a). Every row has the same length, and an entry's position in its row is a segment
// offset of the alpha string plus a rank within a list of the beta string, both
// looked up by enumeration index: 0 the string itself, 1 + i nv + a the single
// occ[i] -> vir[a], 1 + ns + po npv + pv the double occ[i], occ[k] -> vir[a], vir[l]
// with (i < k) = pair po and (a < l) = pair pv (the decode order of k_rows). The
// fill then writes each entry at its place and no sort is needed.
// The tables also give each excitation's colex rank (the column is rank(a') NB +
// rank(b')) and its orbitals, packed i | a << 8 (single) or i | a << 8 | k << 16 | l << 24
// (double, i < k occupied, a < l virtual), so the fill needs no rank computation and no
// index decode beyond one division.
struct SortedRows {
    const int *seg;  // [NA][La]: offset in the row of the segment of alpha excitation e
    const int *pf;   // [NB][LbF]: rank of beta excitation e among b, its singles and doubles
    const int *ps;   // [NB][LbS]: rank of beta excitation e < LbS among b and its singles
    const int *rka, *rkb;           // [NA][La], [NB][LbF]: colex rank of the excited string
    const unsigned *ora, *orb;      // [NA][La], [NB][LbF]: its orbitals, packed
    int La, LbF, LbS;
};

// Hit bitmap (bm, nchunk words per row, or null): the count pass stores each
// chunk's ballot of candidates found in the set, and the fill pass then looks up
// only those lanes and skips chunks without any (the candidates are enumerated in
// the same chunks by both passes).
template <int KIND, bool FILL, bool SORTED>
__global__ void __launch_bounds__(256) k_rows(WfnView w, HamView H, i64 row0, i64 row1, i64 ncol, i64 *cnt_full,
                                              i64 *cnt_low, int *maxlen, const i64 *rowptr, int32_t *cols,
                                              double *vals, double *diag, SortedRows T, u64 *bm, i64 nchunk) {
    static_assert(!SORTED || (FILL && KIND == PYCI_WFN_FULLCI), "sorted fill is for the full FullCI space");
    // GenCI: per wave, the virtual orbital pairs (a < l) of its row, packed a | l << 8
    __shared__ unsigned short genci_vp[KIND == PYCI_WFN_GENCI ? ROWS_PER_BLOCK * GENCI_MAXPV : 1];
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
    // sorted fill: the row's alpha and beta string ranks
    const i64 ra = SORTED ? row / w.nsd : 0, rb = SORTED ? row - ra * w.nsd : 0;
    // One candidate per lane (called by the whole wave): class cls, orbitals ii -> jj
    // (and kk -> ll for the second excitation; for the 1-1 class ii -> jj is alpha,
    // kk -> ll beta), enumeration indices ea / eb in beta list lb for the sorted fill.
    // Forms the determinant, looks it up, evaluates the element and counts or stores it.
    u64 *bmrow = bm ? bm + (row - row0) * nchunk : nullptr;
    i64 chunk = 0;
    auto visit = [&](bool valid, int cls, int ii, int jj, int kk, int ll, int ea, int eb, int lb) {
        i64 col = -1;
        double val = 0.0;
        if (FILL && bmrow) valid = valid && ((bmrow[chunk] >> lane) & 1ULL);
        ++chunk;
        if (valid) {
            if constexpr (SORTED) {
                col = (i64)T.rka[ra * T.La + ea] * w.nsd + (i64)T.rkb[rb * T.LbF + eb];
            } else {
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
            }
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
        if constexpr (SORTED) {
            if (col >= 0) {
                const int pos = T.seg[ra * T.La + ea] +
                                (lb == 1 ? T.ps[rb * T.LbS + eb] : lb == 2 ? T.pf[rb * T.LbF + eb] : 0);
                cols[out_base + pos] = (int32_t)col;
                vals[out_base + pos] = val;
            }
            return;
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
            if (bmrow && lane == 0) bmrow[chunk - 1] = hit;
            cfull += __popcll(hit);
            clow += __popcll(__ballot(col >= 0 && col < row));
        }
    };
    if constexpr (KIND == PYCI_WFN_GENCI) {
        // GenCI: the singles flat, then the doubles with the occupied pair uniform
        // over the wave and the virtual pairs from a per-wave LDS table built once per
        // row (no per-candidate index decode). The order of the candidates is free:
        // the row is sorted by column afterwards.
        // With w.szh > 0 every determinant of the set has the same number of occupied
        // orbitals below szh, so only excitations that keep that number can find a
        // column: singles within each half, doubles whose virtual pair has as many
        // orbitals below szh as the occupied pair (0, 1 or 2). The orbital lists are
        // sorted, so the lower orbitals come first (nol occupied, nvl virtual); the
        // virtual pair table holds the three classes in turn (both low, one low, none),
        // each pair as (a < l) like the unfiltered table, so every visited candidate
        // gets the same orbitals, phase and value. The count and fill passes enumerate
        // the same chunks either way.
        const u64 lowm = w.szh > 0 ? full_mask(w.szh) : 0ULL;
        const int nol = w.szh > 0 ? __popcll(up & lowm) : nou, nvl = w.szh > 0 ? __popcll(~up & m & lowm) : nvu;
        const int noh = nou - nol, nvh = nvu - nvl;
        const int sl = nol * nvl, ns = sl + noh * nvh;
        for (int e0 = 0; e0 < ns; e0 += 64) {
            const int e = e0 + lane, ec = e < ns ? e : 0;
            int i, a;
            if (ec < sl) {
                i = ec / nvl;
                a = ec - i * nvl;
            } else {
                const int t = ec - sl;
                i = nol + t / nvh;
                a = nvl + (t - (i - nol) * nvh);
            }
            const int ii = __shfl(packed, i) & 255, jj = (__shfl(packed, a) >> 8) & 255;
            visit(e < ns, 0, ii, jj, 0, 0, 0, 0, 0);
        }
        const int nLL = nvl * (nvl - 1) / 2, nLH = nvl * nvh, nHH = nvh * (nvh - 1) / 2;
        unsigned short *vp = genci_vp + (threadIdx.x >> 6) * GENCI_MAXPV;
        for (int t0 = 0; t0 < npv_u; t0 += 64) {
            const int t = t0 + lane, tt = t < npv_u ? t : 0;
            int a, l;
            if (tt < nLL) {
                tri_decode(tt, a, l);
            } else if (tt < nLL + nLH) {
                const int u = tt - nLL;
                a = u / nvh;
                l = nvl + (u - a * nvh);
            } else {
                tri_decode(tt - nLL - nLH, a, l);
                a += nvl;
                l += nvl;
            }
            const int va = (__shfl(packed, a) >> 8) & 255, vl = (__shfl(packed, l) >> 8) & 255;
            if (t < npv_u) vp[t] = (unsigned short)(va | (vl << 8));
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        for (int po = 0; po < npo_u; ++po) {
            int i, k;
            tri_decode(po, i, k);
            const int ii = __shfl(packed, i) & 255, kk = __shfl(packed, k) & 255;
            const int nlo = (i < nol) + (k < nol);  // occupied orbitals of the pair below szh
            const int s0 = nlo == 2 ? 0 : nlo == 1 ? nLL : nLL + nLH;
            const int sn = w.szh == 0 ? npv_u : nlo == 2 ? nLL : nlo == 1 ? nLH : nHH;
            for (int t0 = 0; t0 < sn; t0 += 64) {
                const int t = t0 + lane;
                const unsigned pv = vp[(w.szh == 0 ? 0 : s0) + (t < sn ? t : 0)];
                visit(t < sn, 2, ii, (int)(pv & 255), kk, (int)(pv >> 8), 0, 0, 0);
            }
        }
    } else if constexpr (SORTED) {
        // full FullCI space: orbitals and columns from the string tables
        const float rnsd = 1.0f / (float)(nsd > 0 ? nsd : 1);
        for (int e0 = 0; e0 < ncand; e0 += 64) {
            const int e = e0 + lane;
            const bool valid = e < ncand;
            int cls = 0, ea = 0, eb = 0, lb = 0;
            if (valid) {
                if (e < c0) {
                    cls = 0;
                    ea = 1 + e;
                    lb = 1;
                } else if (e < c1) {
                    cls = 1;
                    const int t = e - c0;
                    int ia = (int)((float)t * rnsd);  // t / nsd, corrected below (t < 2^24)
                    if (ia * nsd > t) --ia;
                    else if ((ia + 1) * nsd <= t) ++ia;
                    ea = 1 + ia;
                    eb = 1 + (t - ia * nsd);
                    lb = 1;
                } else if (e < c2) {
                    cls = 2;
                    ea = 1 + nsu + (e - c1);
                } else if (e < c3) {
                    cls = 3;
                    eb = 1 + (e - c2);
                    lb = 2;
                } else {
                    cls = 4;
                    eb = 1 + nsd + (e - c3);
                    lb = 2;
                }
            }
            const unsigned oa = T.ora[ra * T.La + ea], ob = T.orb[rb * T.LbF + eb];
            // alpha orbitals for the classes 0-2, beta for 3-4; the 1-1 class takes both
            const unsigned o1 = cls <= 2 ? oa : ob, o2 = cls == 1 ? (ob << 16) : o1;
            visit(valid, cls, (int)(o1 & 255u), (int)((o1 >> 8) & 255u), (int)((o2 >> 16) & 255u),
                  (int)((o2 >> 24) & 255u), ea, eb, lb);
        }
    } else {
        for (int e0 = 0; e0 < ncand; e0 += 64) {
            const int e = e0 + lane;
            const bool valid = e < ncand;
            int cls = 0, i = 0, a = 0, k = 0, l = 0;
            int ea = 0, eb = 0, lb = 0;  // enumeration indices; beta list: 0 none, 1 singles, 2 all
            if (valid) {
                if (e < c0) {
                    cls = 0;
                    i = e / nvu;
                    a = e - i * nvu;
                    ea = 1 + e;
                    lb = 1;
                } else if (e < c1) {
                    cls = 1;
                    int t = e - c0, ia = t / nsd, kl = t - ia * nsd;
                    i = ia / nvu;
                    a = ia - i * nvu;
                    k = kl / nvd;
                    l = kl - k * nvd;
                    ea = 1 + ia;
                    eb = 1 + kl;
                    lb = 1;
                } else if (e < c2) {
                    cls = 2;
                    int t = e - c1, po = t / npv_u, pv = t - po * npv_u;
                    tri_decode(po, i, k);
                    tri_decode(pv, a, l);
                    ea = 1 + nsu + t;
                } else if (e < c3) {
                    cls = 3;
                    int t = e - c2;
                    i = t / nvd;
                    a = t - i * nvd;
                    eb = 1 + t;
                    lb = 2;
                } else {
                    cls = 4;
                    int t = e - c3, po = t / npv_d, pv = t - po * npv_d;
                    tri_decode(po, i, k);
                    tri_decode(pv, a, l);
                    eb = 1 + nsd + t;
                    lb = 2;
                }
            }
            // gather orbital numbers from the owning lanes (uniform control flow)
            const int pi = __shfl(packed, i), pa = __shfl(packed, a), pk = __shfl(packed, k), pl = __shfl(packed, l);
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
            visit(valid, cls, ii, jj, kk, ll, ea, eb, lb);
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
                const int pos = SORTED ? T.seg[ra * T.La] + T.pf[rb * T.LbF] : written;
                cols[out_base + pos] = (int32_t)row;
                vals[out_base + pos] = d;
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
// Per-row sort by column (sort_row, sparseop.cpp:214-218). Rows of up to 4,096
// entries: a bitonic network in LDS in the all-ascending form, so padding past the
// row end behaves as +infinity, with the LDS sized for CAP entries (rows of at most
// 1,024 entries, most rows of a selected set, in the 12 KB variant, so that more
// blocks are resident). Longer rows: a segmented radix sort over a gathered copy.
// ---------------------------------------------------------------------------
constexpr int SORT_CAP = 4096;
constexpr int SORT_THREADS = 256;

// one block per row; rows with lo < len <= CAP
template <int CAP>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_rows_lds(i64 nrow, const i64 *rowptr, int32_t *cols,
                                                                double *vals, int lo) {
    __shared__ int32_t sk[CAP];
    __shared__ double sv[CAP];
    const i64 row = blockIdx.x;
    if (row >= nrow) return;
    const i64 b = rowptr[row];
    const int len = (int)(rowptr[row + 1] - b);
    if (len <= lo || len <= 1 || len > CAP) return;  // uniform per block
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

// Rows longer than SORT_CAP: their ids (in no particular order) and entry count.
__global__ void k_long_rows(i64 nrow, const i64 *rowptr, unsigned long long *cnt, i64 *ids) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const i64 len = rowptr[r + 1] - rowptr[r];
    if (len <= SORT_CAP) return;
    const unsigned long long o = atomicAdd(cnt, 1ULL);
    atomicAdd(cnt + 1, (unsigned long long)len);
    ids[o] = r;
}

__global__ void k_long_lens(const i64 *ids, i64 L, const i64 *rowptr, i64 *off) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= L) off[i] = (i < L) ? rowptr[ids[i] + 1] - rowptr[ids[i]] : 0;
}

// off[0 .. n] - base (offsets of one batch of long rows, relative to the batch)
__global__ void k_long_rebase(const i64 *off, i64 n, i64 base, i64 *out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = off[i] - base;
}

// one block per long row: to (dir 0) or from (dir 1) the gathered copy
__global__ void k_long_move(const i64 *ids, const i64 *off, const i64 *rowptr, int32_t *cols, double *vals,
                            int32_t *gk, double *gv, int dir) {
    const i64 i = blockIdx.x, b = rowptr[ids[i]], g = off[i], len = off[i + 1] - g;
    for (i64 t = threadIdx.x; t < len; t += blockDim.x) {
        if (dir == 0) {
            gk[g + t] = cols[b + t];
            gv[g + t] = vals[b + t];
        } else {
            cols[b + t] = gk[g + t];
            vals[b + t] = gv[g + t];
        }
    }
}

// Rows longer than SORT_CAP, in batches of at most kLongBatch entries (a row is
// shorter than the 2^31 columns of the operator): gather, segmented radix sort,
// scatter back.
constexpr i64 kLongBatch = (i64)1 << 27;

static void sort_long_rows(pyci_ctx *ctx, i64 nr, const i64 *rowptr, int32_t *cols, double *vals, hipStream_t s) {
    DevBuf<unsigned long long> cnt(2);
    DevBuf<i64> ids((size_t)nr);
    PE_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * 2, s));
    k_long_rows<<<grid_for(nr, 256), 256, 0, s>>>(nr, rowptr, cnt.p, ids.p);
    PE_LAUNCH_CHECK();
    unsigned long long h[2] = {0, 0};
    PE_HIP(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    const i64 L = (i64)h[0];
    if (L == 0) return;
    DevBuf<i64> off((size_t)(L + 1));
    k_long_lens<<<grid_for(L + 1, 256), 256, 0, s>>>(ids.p, L, rowptr, off.p);
    PE_LAUNCH_CHECK();
    exclusive_scan(ctx, off.p, L + 1);
    std::vector<i64> hoff((size_t)(L + 1));
    PE_HIP(hipMemcpyAsync(hoff.data(), off.p, sizeof(i64) * (L + 1), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    i64 cap = 0;  // the largest batch
    for (i64 b = 0; b < L;) {
        i64 e = b + 1;
        while (e < L && hoff[(size_t)(e + 1)] - hoff[(size_t)b] <= kLongBatch) ++e;
        cap = std::max(cap, hoff[(size_t)e] - hoff[(size_t)b]);
        b = e;
    }
    PE_REQUIRE(cap < ((i64)1 << 31), PYCI_EUNSUPPORTED, "row too long to sort");
    DevBuf<int32_t> k0((size_t)cap), k1((size_t)cap);
    DevBuf<double> v0((size_t)cap), v1((size_t)cap);
    DevBuf<i64> loc((size_t)(L + 1));
    DevBuf<unsigned char> tmp;
    for (i64 b = 0; b < L;) {
        i64 e = b + 1;
        while (e < L && hoff[(size_t)(e + 1)] - hoff[(size_t)b] <= kLongBatch) ++e;
        const i64 nrow_b = e - b, M = hoff[(size_t)e] - hoff[(size_t)b];
        k_long_rebase<<<grid_for(nrow_b + 1, 256), 256, 0, s>>>(off.p + b, nrow_b, hoff[(size_t)b], loc.p);
        PE_LAUNCH_CHECK();
        k_long_move<<<(unsigned)nrow_b, 256, 0, s>>>(ids.p + b, loc.p, rowptr, cols, vals, k0.p, v0.p, 0);
        PE_LAUNCH_CHECK();
        size_t tmp_bytes = 0;
        PE_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, v0.p, v1.p, (int)M,
                                                           (int)nrow_b, loc.p, loc.p + 1, 0, 32, s));
        if (tmp.n < tmp_bytes) {
            PE_HIP(hipStreamSynchronize(s));  // the previous batch may still use the old buffer
            tmp.alloc(tmp_bytes);
        }
        PE_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, tmp_bytes, k0.p, k1.p, v0.p, v1.p, (int)M,
                                                           (int)nrow_b, loc.p, loc.p + 1, 0, 32, s));
        k_long_move<<<(unsigned)nrow_b, 256, 0, s>>>(ids.p + b, loc.p, rowptr, cols, vals, k1.p, v1.p, 1);
        PE_LAUNCH_CHECK();
        b = e;
    }
    PE_HIP(hipStreamSynchronize(s));  // the gathered copies leave scope
}

// ---------------------------------------------------------------------------
// Tables of the sorted fill (see SortedRows): one block per string r of nocc
// electrons in n orbitals. Its La = 1 + ns + nd related strings (itself, ns = nocc nv
// singles, nd = C(nocc, 2) C(nv, 2) doubles) are ranked in colex order (bitonic sort
// in LDS); seg (alpha side) receives each one's segment offset in a row, with segment
// sizes LbF (the string itself), LbS (a single) and 1 (a double); pf / ps (beta side)
// its rank among all La / among the string and its singles.
// ---------------------------------------------------------------------------
constexpr int TAB_CAP = 4096;

__global__ void __launch_bounds__(256) k_string_tables(int n, int nocc, const i64 *binom, int LbF, int LbS, int *seg,
                                                       int *pf, int *ps, int *rk, unsigned *orb) {
    __shared__ int key[TAB_CAP];
    __shared__ unsigned short pay[TAB_CAP];
    __shared__ int part[2][256];
    const i64 r = blockIdx.x;
    const int tid = threadIdx.x, nv = n - nocc, ns = nocc * nv, npv = nv * (nv - 1) / 2;
    const int L = 1 + ns + (nocc * (nocc - 1) / 2) * npv;
    // string of colex rank r
    u64 str = 0;
    {
        i64 rank = r;
        int p = n;
        for (int k = nocc; k >= 1; --k) {
            int q = p - 1;
            while (binom[q * 65 + k] > rank) --q;
            str |= 1ULL << q;
            rank -= binom[q * 65 + k];
            p = q;
        }
    }
    const u64 vir = ~str & full_mask(n);
    int N = 1;
    while (N < L) N <<= 1;
    for (int e = tid; e < N; e += 256) {
        if (e < L) {
            u64 t = str;
            unsigned o = 0;
            if (e >= 1 && e <= ns) {
                const int i = (e - 1) / nv, a = (e - 1) - i * nv;
                const int oi = nth_bit(str, i), va = nth_bit(vir, a);
                t ^= (1ULL << oi) | (1ULL << va);
                o = (unsigned)oi | ((unsigned)va << 8);
            } else if (e > ns) {
                const int d = e - 1 - ns, po = d / npv, pv = d - po * npv;
                int i, k, a, l;
                tri_decode(po, i, k);
                tri_decode(pv, a, l);
                const int oi = nth_bit(str, i), ok = nth_bit(str, k), va = nth_bit(vir, a), vl = nth_bit(vir, l);
                t ^= (1ULL << oi) | (1ULL << ok) | (1ULL << va) | (1ULL << vl);
                o = (unsigned)oi | ((unsigned)va << 8) | ((unsigned)ok << 16) | ((unsigned)vl << 24);
            }
            key[e] = (int)colex_rank(t, binom);
            pay[e] = (unsigned short)e;
            rk[r * L + e] = key[e];
            orb[r * L + e] = o;
        } else {
            key[e] = 0x7fffffff;
            pay[e] = 0xffff;
        }
    }
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < N / 2; t += 256) {
                int p, q;
                if (stride == (size >> 1)) {  // first step of a merge: compare mirrored pairs
                    const int blk = t / stride, o = t - blk * stride;
                    p = blk * size + o;
                    q = blk * size + size - 1 - o;
                } else {
                    const int blk = t / stride, o = t - blk * stride;
                    p = blk * 2 * stride + o;
                    q = p + stride;
                }
                if (key[q] < key[p]) {
                    const int tk = key[p];
                    key[p] = key[q];
                    key[q] = tk;
                    const unsigned short tp = pay[p];
                    pay[p] = pay[q];
                    pay[q] = tp;
                }
            }
            __syncthreads();
        }
    }
    // segment offsets and ranks among the string and its singles: block scan over
    // contiguous chunks of the sorted order
    const int CH = (L + 255) / 256, j0 = min(L, tid * CH), j1 = min(L, j0 + CH);
    int sz = 0, c1 = 0;
    for (int j = j0; j < j1; ++j) {
        const int e = pay[j];
        sz += e == 0 ? LbF : e <= ns ? LbS : 1;
        c1 += e <= ns ? 1 : 0;
    }
    part[0][tid] = sz;
    part[1][tid] = c1;
    __syncthreads();
    if (tid == 0) {
        int a = 0, b = 0;
        for (int t = 0; t < 256; ++t) {
            const int x = part[0][t], y = part[1][t];
            part[0][t] = a;
            part[1][t] = b;
            a += x;
            b += y;
        }
    }
    __syncthreads();
    int o = part[0][tid], q = part[1][tid];
    for (int j = j0; j < j1; ++j) {
        const int e = pay[j];
        if (seg) seg[r * L + e] = o;
        if (pf) pf[r * L + e] = j;
        if (ps && e <= ns) ps[r * (1 + ns) + e] = q;
        o += e == 0 ? LbF : e <= ns ? LbS : 1;
        q += e <= ns ? 1 : 0;
    }
}

// Row pointers (already scanned) and per-row lower counts (entries with column <=
// row: the diagonal's position + 1) of rows row0 .. row0 + nr - 1 of the sorted fill.
__global__ void k_sorted_meta(i64 row0, i64 nr, i64 rowlen, i64 nsd, SortedRows T, i64 *rowptr, i64 *lowcnt) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i <= nr; i += (i64)gridDim.x * blockDim.x) {
        rowptr[i] = i * rowlen;
        if (i < nr) {
            const i64 row = row0 + i, ra = row / nsd, rb = row - ra * nsd;
            lowcnt[i] = (i64)T.seg[ra * T.La] + (i64)T.pf[rb * T.LbF] + 1;
        } else {
            lowcnt[i] = 0;
        }
    }
}

// lo[0] = min, lo[1] = max over the determinants of the occupied orbitals below h
__global__ void k_low_range(const u64 *dets, i64 nd, int nkey, int h, int *lo) {
    int mn = 64, mx = 0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += (i64)gridDim.x * blockDim.x) {
        const int c = __popcll(dets[i * nkey] & full_mask(h));
        mn = min(mn, c);
        mx = max(mx, c);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o));
        mx = max(mx, __shfl_xor(mx, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(lo, mn);
        atomicMax(lo + 1, mx);
    }
}

// GenCI: nbasis / 2 if every determinant of the set has the same number of occupied
// spin orbitals in the lower half (a set of one Sz in the alpha-then-beta ordering),
// else 0. An excitation that changes that number then cannot find a column.
static int genci_split(const pyci_wfn *wfn, hipStream_t s) {
    const int n = (int)wfn->nbasis;
    if (wfn->kind != PYCI_WFN_GENCI || n < 2 || n % 2 != 0 || wfn->ndet < 1) return 0;
    DevBuf<int> lo(2);
    const int init[2] = {64, 0};
    PE_HIP(hipMemcpyAsync(lo.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    k_low_range<<<(unsigned)std::min<i64>(1024, (wfn->ndet + 255) / 256), 256, 0, s>>>(wfn->dets.p, wfn->ndet,
                                                                                     (int)wfn->nkey, n / 2, lo.p);
    PE_LAUNCH_CHECK();
    int h[2] = {0, 1};
    PE_HIP(hipMemcpyAsync(h, lo.p, sizeof(h), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    return h[0] == h[1] ? n / 2 : 0;
}

template <int KIND>
static void launch_rows(bool fill, const WfnView &w, const HamView &H, i64 row0, i64 row1, i64 ncol, i64 *cf,
                        i64 *cl, int *mx, const i64 *rp, int32_t *cols, double *vals, double *diag, hipStream_t s,
                        const SortedRows *T = nullptr, u64 *bm = nullptr, i64 nchunk = 0) {
    unsigned g = grid_for(row1 - row0, ROWS_PER_BLOCK);
    if constexpr (KIND == PYCI_WFN_FULLCI) {
        if (T) {
            k_rows<KIND, true, true><<<g, 256, 0, s>>>(w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, *T,
                                                       nullptr, 0);
            PE_LAUNCH_CHECK();
            return;
        }
    }
    const SortedRows none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0};
    if (fill)
        k_rows<KIND, true, false><<<g, 256, 0, s>>>(w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, none,
                                                    bm, nchunk);
    else
        k_rows<KIND, false, false><<<g, 256, 0, s>>>(w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, none,
                                                     bm, nchunk);
    PE_LAUNCH_CHECK();
}

static void dispatch_rows(int kind, bool fill, const WfnView &w, const HamView &H, i64 row0, i64 row1, i64 ncol,
                          i64 *cf, i64 *cl, int *mx, const i64 *rp, int32_t *cols, double *vals, double *diag,
                          hipStream_t s, u64 *bm, i64 nchunk) {
    if (row1 <= row0) return;
    if (kind == PYCI_WFN_DOCI)
        launch_rows<PYCI_WFN_DOCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s, nullptr, bm,
                                   nchunk);
    else if (kind == PYCI_WFN_FULLCI)
        launch_rows<PYCI_WFN_FULLCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s, nullptr, bm,
                                     nchunk);
    else
        launch_rows<PYCI_WFN_GENCI>(fill, w, H, row0, row1, ncol, cf, cl, mx, rp, cols, vals, diag, s, nullptr, bm,
                                    nchunk);
}

// Chunks of 64 candidates k_rows visits per row (the same for every row of a set).
static i64 row_chunks(int kind, i64 n, i64 nu, i64 nd) {
    const i64 vu = n - nu, vd = n - nd, su = nu * vu, sd = nd * vd;
    const i64 pou = nu * (nu - 1) / 2, pvu = vu * (vu - 1) / 2, pod = nd * (nd - 1) / 2, pvd = vd * (vd - 1) / 2;
    auto ch = [](i64 x) { return (x + 63) / 64; };
    if (kind == PYCI_WFN_GENCI) return ch(su) + pou * ch(pvu);
    if (kind == PYCI_WFN_DOCI) return ch(su);
    return ch(su + su * sd + pou * pvu + sd + pod * pvd);
}

// Sorted fill of rows row0 .. row1 - 1 of the full FullCI space (see SortedRows):
// tables, row pointers, lower counts, one fill pass. False where the space or the
// column range is not the full one, or a string has more related strings than the
// table kernel holds.
static bool build_rows_sorted(pyci_op *op, const pyci_wfn *wfn, const WfnView &w, const HamView &H, i64 row0,
                              i64 row1, i64 ncol, DevBuf<i64> &rowptr_seg, DevBuf<int32_t> &cols,
                              DevBuf<double> &vals, DevBuf<i64> &lowcnt_seg, DevBuf<double> &diag_seg, i64 &nnz_seg,
                              i64 &low_total) {
    if (op->kind != PYCI_WFN_FULLCI || !wfn->colex || ncol != wfn->ndet) return false;
    const i64 n = wfn->nbasis, nu = wfn->nup, nd = wfn->ndn, vu = n - nu, vd = n - nd;
    const i64 La = 1 + nu * vu + (nu * (nu - 1) / 2) * (vu * (vu - 1) / 2);
    const i64 LbF = 1 + nd * vd + (nd * (nd - 1) / 2) * (vd * (vd - 1) / 2), LbS = 1 + nd * vd;
    if (La > TAB_CAP || LbF > TAB_CAP) return false;
    const i64 NA = binom_host(n, nu), NB = binom_host(n, nd);
    const i64 rowlen = LbF + nu * vu * LbS + (La - 1 - nu * vu);
    const i64 nr = row1 - row0;
    PE_REQUIRE(rowlen * nr < (i64)1 << 40, PYCI_ENOMEM, "sparse operator too large");
    pyci_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    DevBuf<int> seg((size_t)(NA * La)), pf((size_t)(NB * LbF)), ps((size_t)(NB * LbS));
    DevBuf<int> rka((size_t)(NA * La)), rkb((size_t)(NB * LbF));
    DevBuf<unsigned> ora((size_t)(NA * La)), orb((size_t)(NB * LbF));
    k_string_tables<<<(unsigned)NA, 256, 0, s>>>((int)n, (int)nu, w.binom, (int)LbF, (int)LbS, seg.p, nullptr,
                                                 nullptr, rka.p, ora.p);
    PE_LAUNCH_CHECK();
    k_string_tables<<<(unsigned)NB, 256, 0, s>>>((int)n, (int)nd, w.binom, (int)LbF, (int)LbS, nullptr, pf.p, ps.p,
                                                 rkb.p, orb.p);
    PE_LAUNCH_CHECK();
    const SortedRows T{seg.p, pf.p, ps.p, rka.p, rkb.p, ora.p, orb.p, (int)La, (int)LbF, (int)LbS};
    rowptr_seg.alloc((size_t)(nr + 1));
    lowcnt_seg.alloc((size_t)(nr + 1));
    diag_seg.alloc((size_t)(nr > 0 ? nr : 1));
    k_sorted_meta<<<std::min<unsigned>(4096u, grid_for(nr + 1, 256)), 256, 0, s>>>(row0, nr, rowlen, w.nsd, T,
                                                                                  rowptr_seg.p, lowcnt_seg.p);
    PE_LAUNCH_CHECK();
    {
        DevBuf<i64> tmp((size_t)(nr + 1));
        PE_HIP(hipMemcpyAsync(tmp.p, lowcnt_seg.p, sizeof(i64) * (nr + 1), hipMemcpyDeviceToDevice, s));
        low_total = exclusive_scan(ctx, tmp.p, nr + 1);
    }
    nnz_seg = rowlen * nr;
    cols.alloc((size_t)(nnz_seg > 0 ? nnz_seg : 1));
    vals.alloc((size_t)(nnz_seg > 0 ? nnz_seg : 1));
    if (nr > 0)
        launch_rows<PYCI_WFN_FULLCI>(true, w, H, row0, row1, ncol, nullptr, nullptr, nullptr, rowptr_seg.p, cols.p,
                                     vals.p, diag_seg.p, s, &T);
    PE_HIP(hipStreamSynchronize(s));  // the tables are freed on return
    return true;
}

void build_rows(pyci_op *op, const pyci_ham *ham, const pyci_wfn *wfn, i64 row0, i64 row1, i64 ncol,
                DevBuf<i64> &rowptr_seg, DevBuf<int32_t> &cols, DevBuf<double> &vals, DevBuf<i64> &lowcnt_seg,
                DevBuf<double> &diag_seg, i64 &nnz_seg, i64 &low_total) {
    pyci_ctx *ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const i64 nr = row1 - row0;
    WfnView w = view_of(wfn);
    HamView H = view_of(ham);
    if (build_rows_sorted(op, wfn, w, H, row0, row1, ncol, rowptr_seg, cols, vals, lowcnt_seg, diag_seg, nnz_seg,
                          low_total))
        return;
    w.szh = genci_split(wfn, s);
    rowptr_seg.alloc((size_t)(nr + 1));
    lowcnt_seg.alloc((size_t)(nr + 1));
    diag_seg.alloc((size_t)(nr > 0 ? nr : 1));
    DevBuf<int> mx(1);
    PE_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), s));
    PE_HIP(hipMemsetAsync(rowptr_seg.p, 0, sizeof(i64) * (nr + 1), s));
    PE_HIP(hipMemsetAsync(lowcnt_seg.p, 0, sizeof(i64) * (nr + 1), s));
    // the hit bitmap where it fits in a third of the free memory (36.6 GB at the 6.9 M
    // determinants of the GenCI selection workload); should the CSR arrays then not
    // fit beside it, it is released and the fill pass enumerates every candidate again
    const i64 nchunk = row_chunks(op->kind, wfn->nbasis, wfn->nup, op->kind == PYCI_WFN_FULLCI ? wfn->ndn : 0);
    DevBuf<u64> bm;
    {
        size_t free_b = 0, total_b = 0;
        PE_HIP(hipMemGetInfo(&free_b, &total_b));
        const long double need = (long double)nr * (long double)nchunk * 8.0L;
        if (nr > 0 && need <= (long double)free_b / 3.0L) bm.alloc((size_t)(nr * nchunk));
    }
    dispatch_rows(op->kind, false, w, H, row0, row1, ncol, rowptr_seg.p, lowcnt_seg.p, mx.p, nullptr, nullptr,
                  nullptr, nullptr, s, bm.p, nchunk);
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
    try {
        cols.alloc((size_t)nnz_seg);
        vals.alloc((size_t)nnz_seg);
    } catch (const EngineError &e) {
        if (e.code != PYCI_ENOMEM || !bm.p) throw;
        (void)hipGetLastError();  // the failed allocation must not be reported by a later launch check
        bm.release();
        cols.alloc((size_t)nnz_seg);
        vals.alloc((size_t)nnz_seg);
    }
    dispatch_rows(op->kind, true, w, H, row0, row1, ncol, nullptr, nullptr, nullptr, rowptr_seg.p, cols.p, vals.p,
                  diag_seg.p, s, bm.p, nchunk);
    if (nr > 0) {
        k_sort_rows_lds<1024><<<(unsigned)nr, SORT_THREADS, 0, s>>>(nr, rowptr_seg.p, cols.p, vals.p, 1);
        PE_LAUNCH_CHECK();
        if (hmax > 1024) {
            k_sort_rows_lds<SORT_CAP><<<(unsigned)nr, SORT_THREADS, 0, s>>>(nr, rowptr_seg.p, cols.p, vals.p, 1024);
            PE_LAUNCH_CHECK();
        }
        if (hmax > SORT_CAP) sort_long_rows(ctx, nr, rowptr_seg.p, cols.p, vals.p, s);
    }
    if (bm.p) PE_HIP(hipStreamSynchronize(s));  // the fill pass has read the bitmap before it is freed
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
