// Heat-bath CI selection on the GPU.
//
// Reference: add_hci (hci.cpp:252-279) with the per-determinant screening of
// hci.cpp:49-238. A candidate D' of source determinant D_i is added when the
// magnitude of its matrix element exceeds eps/|c_i| and D' is not yet in the
// wave function (singles use the full single-excitation element, 1-1 doubles
// |<ik|ab>|, same-spin doubles |<ik|ab> - <ik|ba>|, DOCI pair hops |v_kl|).
//
// The reference merges per-thread hash maps, so its append order depends on the
// hash table's iteration order. Here new determinants are appended in the order
// of their FIRST occurrence in the serial enumeration (source det ascending,
// then the reference's loop order inside a row), which is deterministic:
//   pass 1 counts passing new candidates (with repeats) to size a hash set,
//   pass 2 inserts each candidate's 64-bit key with the minimum position,
//   then the claimed slots, one per new determinant, are sorted by that
//   position (a radix sort) and written in that order.
// One wavefront handles one source row; the outer (i, j) loops run uniformly
// and lanes cover the inner excitation block, so ballot compaction preserves
// the reference order.
//
// The same passes also build a candidate pool for the learned selector: the
// insertion threshold is then a looser eps_pool, and each candidate records its
// heat-bath score max_i |H_ji c_i| and whether plain add_hci(eps) takes it.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace pe {

constexpr u64 EMPTY_KEY = ~0ULL;

struct NewSet {
    u64 *keys;
    u64 *pos;
    u64 mask;
    unsigned long long *unique;  // number of claimed slots
    int *overflow;
    u64 *score = nullptr;  // pool only: max |H_ji c_i| as double bits (nonnegative, so integer max)
    int *plain = nullptr;  // pool only: 1 when some source passes the plain threshold
    double *num = nullptr; // ENPT2 only: sum_i H_ai c_i of external determinant a
    u64 *kdet = nullptr;   // ENPT2 only: the determinant words of each claimed slot (2 per slot)
};

__device__ inline u64 set_key(const WfnView &w, u64 a, u64 b) {
    if (w.nkey == 2) return (u64)colex_rank(a, w.binom) * (u64)w.nsd + (u64)colex_rank(b, w.binom);
    return a;
}

// Inserts key with position p (keeps the minimum); returns the slot, or ~0 on
// overflow. Claimed slots are counted.
__device__ inline u64 set_insert(const NewSet &S, u64 key, u64 p) {
    u64 h = mix64(key) & S.mask;
    for (u64 step = 0; step <= S.mask; ++step) {
        unsigned long long old = atomicCAS((unsigned long long *)&S.keys[h], (unsigned long long)EMPTY_KEY,
                                           (unsigned long long)key);
        if (old == EMPTY_KEY) atomicAdd(S.unique, 1ULL);
        if (old == EMPTY_KEY || old == key) {
            atomicMin((unsigned long long *)&S.pos[h], (unsigned long long)p);
            return h;
        }
        h = (h + 1) & S.mask;
    }
    atomicOr(S.overflow, 1);
    return ~0ULL;
}

#define GIX(p, q, r, s) ((((p) * n + (q)) * n + (r)) * n + (s))

// Heat-bath lists of a GenCI Hamiltonian (n <= 64 spin-orbitals). For every
// orbital pair p < q (pair index P = q (q - 1) / 2 + p) the pairs r < s ordered by
// |V| descending, V = two_mo[p][q][r][s] - two_mo[p][q][s][r] (the double-excitation
// element of add_hci, hci.cpp:219-220, with the same operands, so the same bits),
// stored with its sign; for every (i, a) an upper bound of the single-excitation
// element: |h_ia| + sum over all q of |two_mo[i][q][a][q] - two_mo[i][q][q][a]|.
// A row then visits, per occupied pair, only the list prefix above eps / |c_i|, and
// computes a single's element only where the bound reaches it.
struct HbLists {
    const double *v = nullptr;         // [P][npair]
    const unsigned short *rs = nullptr;  // [P][npair]: r | s << 8
    const double *sb = nullptr;        // [i][a]
    int npair = 0;
};
constexpr int HB_CAP = 2048;  // n (n - 1) / 2 <= 2016 for n <= 64

__global__ void __launch_bounds__(256) k_hb_lists(HamView H, int npair, double *lv, unsigned short *lrs) {
    __shared__ double key[HB_CAP];
    __shared__ unsigned short idx[HB_CAP];
    const int n = H.n, tid = threadIdx.x;
    const double *two = H.two;
    int p, q;
    tri_decode((int)blockIdx.x, p, q);
    int N = 1;
    while (N < npair) N <<= 1;
    for (int j = tid; j < N; j += 256) {
        if (j < npair) {
            int r, s;
            tri_decode(j, r, s);
            key[j] = fabs(two[GIX(p, q, r, s)] - two[GIX(p, q, s, r)]);
        } else {
            key[j] = -1.0;
        }
        idx[j] = (unsigned short)j;
    }
    __syncthreads();
    // bitonic sort, descending by |V|
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < N / 2; t += 256) {
                const int blk = t / stride, o = t - blk * stride;
                int a, b;
                if (stride == (size >> 1)) {
                    a = blk * size + o;
                    b = blk * size + size - 1 - o;
                } else {
                    a = blk * 2 * stride + o;
                    b = a + stride;
                }
                if (key[a] < key[b]) {
                    const double tk = key[a];
                    key[a] = key[b];
                    key[b] = tk;
                    const unsigned short ti = idx[a];
                    idx[a] = idx[b];
                    idx[b] = ti;
                }
            }
            __syncthreads();
        }
    }
    const i64 base = (i64)blockIdx.x * npair;
    for (int t = tid; t < npair; t += 256) {
        int r, s;
        tri_decode(idx[t], r, s);
        lv[base + t] = two[GIX(p, q, r, s)] - two[GIX(p, q, s, r)];
        lrs[base + t] = (unsigned short)(r | (s << 8));
    }
}

__global__ void k_hb_single_bound(HamView H, double *sb) {
    const int n = H.n, e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n * n) return;
    const double *one = H.one, *two = H.two;
    const int i = e / n, a = e - i * n;
    double b = fabs(one[i * n + a]);
    for (int q = 0; q < n; ++q) b += fabs(two[GIX(i, q, a, q)] - two[GIX(i, q, q, a)]);
    sb[e] = b;
}

// PASS 1: count, 2: insert, 5: ENPT2 numerators (insert, then add the signed
// H_ai c_i of every passing candidate).
template <int KIND, int PASS>
__global__ void __launch_bounds__(256) k_hci(WfnView w, HamView H, const double *coeffs, double eps, double eps_plain,
                                             i64 rows_stride, NewSet S, unsigned long long *count, HbLists L) {
    const int lane = threadIdx.x & 63;
    const i64 row = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= w.ndet) return;
    const int n = H.n;
    const double *one = H.one, *two = H.two;
    const u64 m = full_mask(n);
    const u64 up = w.dets[(i64)w.nkey * row];
    const u64 dn = (KIND == PYCI_WFN_FULLCI) ? w.dets[2 * row + 1] : 0ULL;
    const int nou = __popcll(up), nvu = n - nou, nod = __popcll(dn), nvd = n - nod;
    const double cabs = fabs(coeffs[row]);
    const double eps_i = eps / cabs, eps_plain_i = eps_plain / cabs;
    // lane q holds the q-th occupied / virtual orbital of each string
    int tmp;
    auto nth = [](u64 x, int t) {
        for (int s = 0; s < t && x; ++s) x &= x - 1;
        return x ? __builtin_ctzll(x) : 0;
    };
    const int packed = nth(up, lane) | (nth(~up & m, lane) << 8) | (nth(dn, lane) << 16) | (nth(~dn & m, lane) << 24);
    (void)tmp;
    u64 local = 0;                 // PASS 1 count
    i64 e_base = 0;                // running position inside the row (reference order)

    // Evaluates one candidate; returns true when it passes and is new.
    auto consider = [&](bool valid, u64 nu, u64 nd, double val, u64 pos, int sg) {
        bool pass = false;
        u64 key = 0;
        if (valid && fabs(val) > eps_i) {
            if (wfn_lookup(w, nu, nd) < 0) {
                pass = true;
                key = set_key(w, nu, nd);
            }
        }
        if (PASS == 1) {
            local += __popcll(__ballot(pass));
        } else if (PASS == 5) {
            // the reference multiplies the element by (phase * c_i) (enpt2.cpp:122-205)
            if (pass) {
                const u64 slot = set_insert(S, key, 0);
                if (slot != ~0ULL) {
                    atomicAdd(&S.num[slot], val * ((double)sg * coeffs[row]));
                    S.kdet[2 * slot] = nu;
                    S.kdet[2 * slot + 1] = nd;
                }
            }
        } else if (PASS == 2) {
            if (pass) {
                const u64 slot = set_insert(S, key, pos);
                if (S.score != nullptr && slot != ~0ULL) {
                    atomicMax((unsigned long long *)&S.score[slot],
                              (unsigned long long)__double_as_longlong(fabs(val) * cabs));
                    if (fabs(val) > eps_plain_i) atomicOr(&S.plain[slot], 1);
                }
            }
        }
    };

    if constexpr (KIND == PYCI_WFN_GENCI) {
        // Position of a candidate in the reference's row order (for i, for j: single
        // (i, j), then doubles (k > i, l > j)): the single (i, j) at off(i, j), the
        // double (i, j, k, l) at off(i, j) + 1 + (k - i - 1)(nv - 1 - j) + (l - j - 1).
        const int no = nou, nv = nvu;
        const i64 hv = (i64)nv * (nv - 1) / 2;
        auto off = [&](int i, int j) -> i64 {
            return (i64)i * nv + hv * ((i64)i * (no - 1) - (i64)i * (i - 1) / 2) + j +
                   (i64)(no - 1 - i) * ((i64)j * (nv - 1) - (i64)j * (j - 1) / 2);
        };
        const u64 rbase = (u64)row * (u64)rows_stride;
        // singles, where the bound reaches the threshold
        for (int e0 = 0; e0 < no * nv; e0 += 64) {
            const int e = e0 + lane;
            const bool in = e < no * nv;
            const int ec = in ? e : 0, i = ec / nv, j = ec - i * nv;
            const int ii = __shfl(packed, i) & 255, jj = (__shfl(packed, j) >> 8) & 255;
            const bool live = in && L.sb[ii * n + jj] * (1.0 + 1e-12) > eps_i;
            double val = 0.0;
            int sg = 1;
            if (live) {
                double v = one[ii * n + jj];
                for (u64 x = up; x; x &= x - 1) {
                    int q = __builtin_ctzll(x);
                    v += two[GIX(ii, q, jj, q)] - two[GIX(ii, q, q, jj)];
                }
                val = v;
                if (PASS == 5) sg = phase1(up, ii, jj);
            }
            consider(live, up ^ (1ULL << ii) ^ (1ULL << jj), 0ULL, val, rbase + (u64)off(i, j), sg);
        }
        // doubles: per occupied pair, the list prefix above the threshold
        for (int po = 0; po < no * (no - 1) / 2; ++po) {
            int i, k;
            tri_decode(po, i, k);
            const int ii = __shfl(packed, i) & 255, kk = __shfl(packed, k) & 255;
            const i64 lb = (i64)(kk * (kk - 1) / 2 + ii) * L.npair;
            for (int t0 = 0; t0 < L.npair; t0 += 64) {
                const int t = t0 + lane;
                const bool in = t < L.npair;
                const double V = in ? L.v[lb + t] : 0.0;
                const bool above = in && fabs(V) > eps_i;
                if (__ballot(above) == 0) break;  // the list is sorted: nothing further passes
                const unsigned rs = in ? L.rs[lb + t] : 0u;
                const int r = (int)(rs & 255u), s2 = (int)(rs >> 8);
                const bool free = above && !((up >> r) & 1ULL) && !((up >> s2) & 1ULL);
                // r and s as indices among the virtual orbitals
                const int j = r - __popcll(up & ((1ULL << r) - 1ULL)), l = s2 - __popcll(up & ((1ULL << s2) - 1ULL));
                int sg = 1;
                if (PASS == 5 && free) sg = phase2(up, ii, kk, r, s2);
                const u64 pos = rbase + (u64)(off(i, j) + 1 + (i64)(k - i - 1) * (nv - 1 - j) + (l - j - 1));
                consider(free, up ^ (1ULL << ii) ^ (1ULL << kk) ^ (1ULL << r) ^ (1ULL << s2), 0ULL, V, pos, sg);
            }
        }
    } else if (KIND == PYCI_WFN_DOCI) {
        const int total = nou * nvu;
        for (int e0 = 0; e0 < total; e0 += 64) {
            const int e = e0 + lane;
            const bool valid = e < total;
            const int i = valid ? e / nvu : 0, j = valid ? e - (e / nvu) * nvu : 0;
            const int pi = __shfl(packed, i), pj = __shfl(packed, j);
            const int k = pi & 255, l = (pj >> 8) & 255;
            double val = valid ? two[GIX(k, k, l, l)] : 0.0;
            consider(valid, up ^ (1ULL << k) ^ (1ULL << l), 0ULL, val, (u64)row * (u64)rows_stride + (u64)e, 1);
        }
    } else {
        // alpha (or only) string: for i, for j: single, [1-1 block], same-spin doubles k > i, l > j
        for (int i = 0; i < nou; ++i) {
            for (int j = 0; j < nvu; ++j) {
                const int ii = __shfl(packed, i) & 255, jj = (__shfl(packed, j) >> 8) & 255;
                const int nsd = (KIND == PYCI_WFN_FULLCI) ? nod * nvd : 0;
                const int ndbl = (nou - 1 - i) * (nvu - 1 - j);
                const int bsize = 1 + nsd + ndbl;
                for (int t0 = 0; t0 < bsize; t0 += 64) {
                    const int t = t0 + lane;
                    const bool valid = t < bsize;
                    int k = 0, l = 0, cls = 0;
                    if (valid && t > 0) {
                        if (t < 1 + nsd) {
                            cls = 1;
                            k = (t - 1) / nvd;
                            l = (t - 1) - k * nvd;
                        } else {
                            cls = 2;
                            const int u = t - 1 - nsd, w2 = nvu - 1 - j;
                            k = i + 1 + u / w2;
                            l = j + 1 + (u - (u / w2) * w2);
                        }
                    }
                    const int pk = __shfl(packed, k), pl = __shfl(packed, l);
                    u64 nu = up ^ (1ULL << ii) ^ (1ULL << jj), nd = dn;
                    double val = 0.0;
                    int sg = 1;
                    if (valid) {
                        if (PASS == 5) sg = phase1(up, ii, jj);
                        if (cls == 0) {
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
                            val = v;
                        } else if (cls == 1) {
                            const int kk = (pk >> 16) & 255, ll = (pl >> 24) & 255;
                            nd = dn ^ (1ULL << kk) ^ (1ULL << ll);
                            val = two[GIX(ii, kk, jj, ll)];
                            if (PASS == 5) sg *= phase1(dn, kk, ll);
                        } else {
                            const int kk = pk & 255, ll = (pl >> 8) & 255;
                            nu ^= (1ULL << kk) ^ (1ULL << ll);
                            val = two[GIX(ii, kk, jj, ll)] - two[GIX(ii, kk, ll, jj)];
                            if (PASS == 5) sg = phase2(up, ii, kk, jj, ll);
                        }
                    }
                    consider(valid, nu, nd, val, (u64)row * (u64)rows_stride + (u64)(e_base + t), sg);
                }
                e_base += bsize;
            }
        }
        if (KIND == PYCI_WFN_FULLCI) {
            for (int i = 0; i < nod; ++i) {
                for (int j = 0; j < nvd; ++j) {
                    const int ii = (__shfl(packed, i) >> 16) & 255, jj = (__shfl(packed, j) >> 24) & 255;
                    const int ndbl = (nod - 1 - i) * (nvd - 1 - j);
                    const int bsize = 1 + ndbl;
                    for (int t0 = 0; t0 < bsize; t0 += 64) {
                        const int t = t0 + lane;
                        const bool valid = t < bsize;
                        int k = 0, l = 0;
                        if (valid && t > 0) {
                            const int u = t - 1, w2 = nvd - 1 - j;
                            k = i + 1 + u / w2;
                            l = j + 1 + (u - (u / w2) * w2);
                        }
                        const int pk = __shfl(packed, k), pl = __shfl(packed, l);
                        u64 nd = dn ^ (1ULL << ii) ^ (1ULL << jj);
                        double val = 0.0;
                        int sg = 1;
                        if (valid) {
                            if (PASS == 5) sg = phase1(dn, ii, jj);
                            if (t == 0) {
                                double v = one[ii * n + jj];
                                for (u64 x = up; x; x &= x - 1) {
                                    int q = __builtin_ctzll(x);
                                    v += two[GIX(ii, q, jj, q)];
                                }
                                for (u64 x = dn; x; x &= x - 1) {
                                    int q = __builtin_ctzll(x);
                                    v += two[GIX(ii, q, jj, q)] - two[GIX(ii, q, q, jj)];
                                }
                                val = v;
                            } else {
                                const int kk = (pk >> 16) & 255, ll = (pl >> 24) & 255;
                                nd ^= (1ULL << kk) ^ (1ULL << ll);
                                val = two[GIX(ii, kk, jj, ll)] - two[GIX(ii, kk, ll, jj)];
                                if (PASS == 5) sg = phase2(dn, ii, kk, jj, ll);
                            }
                        }
                        consider(valid, up, nd, val, (u64)row * (u64)rows_stride + (u64)(e_base + t), sg);
                    }
                    e_base += bsize;
                }
            }
        }
    }
    if (lane == 0 && PASS == 1) atomicAdd(count, (unsigned long long)local);
}

__global__ void k_set_clear(u64 *keys, u64 *pos, u64 *score, int *plain, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        keys[i] = EMPTY_KEY;
        pos[i] = ~0ULL;
        if (score) {
            score[i] = 0ULL;
            plain[i] = 0;
        }
    }
}

// The heat-bath lists of a GenCI selection (HbLists), built for one call.
struct HbTables {
    DevBuf<double> v, sb;
    DevBuf<unsigned short> rs;
    HbLists L;
    void build(const HamView &H, hipStream_t s) {
        const int n = H.n, npair = n * (n - 1) / 2;
        PE_REQUIRE(npair <= HB_CAP, PYCI_EUNSUPPORTED, "too many orbitals for the heat-bath lists");
        v.alloc((size_t)std::max(1, npair * npair));
        rs.alloc((size_t)std::max(1, npair * npair));
        sb.alloc((size_t)std::max(1, n * n));
        if (npair > 0) {
            k_hb_lists<<<(unsigned)npair, 256, 0, s>>>(H, npair, v.p, rs.p);
            PE_LAUNCH_CHECK();
        }
        k_hb_single_bound<<<grid_for((i64)n * n, 256), 256, 0, s>>>(H, sb.p);
        PE_LAUNCH_CHECK();
        L = HbLists{v.p, rs.p, sb.p, npair};
    }
};

template <int PASS>
static void launch_hci(int kind, const WfnView &w, const HamView &H, const double *c, double eps, double eps_plain,
                       i64 stride, const NewSet &S, unsigned long long *count, hipStream_t s, const HbLists &L) {
    unsigned g = grid_for(w.ndet, 4);
    if (kind == PYCI_WFN_DOCI)
        k_hci<PYCI_WFN_DOCI, PASS><<<g, 256, 0, s>>>(w, H, c, eps, eps_plain, stride, S, count, L);
    else if (kind == PYCI_WFN_FULLCI)
        k_hci<PYCI_WFN_FULLCI, PASS><<<g, 256, 0, s>>>(w, H, c, eps, eps_plain, stride, S, count, L);
    else
        k_hci<PYCI_WFN_GENCI, PASS><<<g, 256, 0, s>>>(w, H, c, eps, eps_plain, stride, S, count, L);
    PE_LAUNCH_CHECK();
}

void wfn_index_append(pyci_wfn *w, i64 first);

// Claimed slots of the set as (first position, slot) pairs, in no particular order:
// one cursor increment per wave.
__global__ void k_set_owners(NewSet S, i64 cap, unsigned long long *cursor, u64 *pos_out, u64 *slot_out) {
    const int lane = threadIdx.x & 63;
    const i64 step = (i64)gridDim.x * blockDim.x;
    for (i64 h0 = (i64)blockIdx.x * blockDim.x + (threadIdx.x & ~63); h0 < cap; h0 += step) {
        const i64 h = h0 + lane;
        const bool claimed = h < cap && S.keys[h] != EMPTY_KEY;
        const u64 b = __ballot(claimed);
        unsigned long long base = 0;
        if (lane == 0 && b) base = atomicAdd(cursor, (unsigned long long)__popcll(b));
        base = __shfl(base, 0);
        if (claimed) {
            const unsigned long long o = base + (unsigned long long)__popcll(b & ((1ULL << lane) - 1ULL));
            pos_out[o] = S.pos[h];
            slot_out[o] = (u64)h;
        }
    }
}

// String of colex rank r with k electrons in nb orbitals.
__device__ inline u64 colex_unrank(i64 r, int nb, int k, const i64 *binom) {
    u64 str = 0;
    int p = nb;
    for (int q = k; q >= 1; --q) {
        int t = p - 1;
        while (binom[t * 65 + q] > r) --t;
        str |= 1ULL << t;
        r -= binom[t * 65 + q];
        p = t;
    }
    return str;
}

// The determinants of the sorted owner slots: a one-word key is the determinant,
// a FullCI key is rank(alpha) * nsd + rank(beta) (set_key).
__global__ void k_emit_owners(WfnView w, NewSet S, const u64 *slots, i64 n, u64 *out) {
    const i64 o = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const u64 key = S.keys[slots[o]];
    if (w.nkey == 1) {
        out[o] = key;
    } else {
        const i64 ra = (i64)(key / (u64)w.nsd), rb = (i64)(key - (u64)ra * (u64)w.nsd);
        out[2 * o] = colex_unrank(ra, w.nb, w.nup, w.binom);
        out[2 * o + 1] = colex_unrank(rb, w.nb, w.ndn, w.binom);
    }
}

// Passes 1-2 of the selection: the set of new candidates (|H_ji| > eps/|c_i|,
// not in the wave function) with their first-occurrence positions; then its
// claimed slots ordered by that position (a radix sort), which is the order of
// first occurrence in the serial enumeration: slots[0 .. nnew).
struct HciRun {
    DevBuf<u64> keys, pos, score, slots;
    DevBuf<int> plain, ovf;
    DevBuf<unsigned long long> cnt;
    NewSet S{};
    i64 stride = 0, nnew = 0;
    unsigned long long S_unique_host = 0;  // claimed slots after pass 2
};

static void hci_collect(pyci_ham *ham, pyci_wfn *wfn, const double *dc, double eps, double eps_plain, bool pool,
                        HciRun &R) {
    pyci_ctx *ctx = wfn->ctx;
    hipStream_t s = ctx->stream;
    const i64 nd = wfn->ndet;
    PE_REQUIRE(ham->nbasis == wfn->nbasis, PYCI_EINVAL, "operator and wave function have different nbasis");
    if (wfn->kind == PYCI_WFN_FULLCI) {
        long double tot = (long double)binom_host(wfn->nbasis, wfn->nup) * (long double)binom_host(wfn->nbasis, wfn->ndn);
        PE_REQUIRE(tot < 9.0e18L, PYCI_EUNSUPPORTED, "determinant space too large for the GPU selection keys");
    }
    // positions: row * stride + in-row index; stride >= candidates per row
    const i64 no = wfn->nup, nv = wfn->nbasis - wfn->nup, ndo = wfn->ndn, ndv = wfn->nbasis - wfn->ndn;
    R.stride = no * nv + no * nv * ndo * ndv + (no * (no - 1) / 2) * (nv * (nv - 1) / 2) + ndo * ndv +
               (ndo * (ndo - 1) / 2) * (ndv * (ndv - 1) / 2) + 1;
    PE_REQUIRE((long double)R.stride * (long double)nd < 1.8e19L, PYCI_EUNSUPPORTED, "selection too large");
    const WfnView w = view_of(wfn);
    const HamView H = view_of(ham);
    R.cnt.alloc(2);
    R.ovf.alloc(1);
    PE_HIP(hipMemsetAsync(R.cnt.p, 0, sizeof(unsigned long long) * 2, s));
    R.S = NewSet{nullptr, nullptr, 0, R.cnt.p + 1, R.ovf.p};
    HbTables hb;
    if (wfn->kind == PYCI_WFN_GENCI) hb.build(H, s);
    launch_hci<1>(wfn->kind, w, H, dc, eps, eps_plain, R.stride, R.S, R.cnt.p, s, hb.L);
    unsigned long long upper = 0;
    PE_HIP(hipMemcpyAsync(&upper, R.cnt.p, sizeof(upper), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    R.nnew = 0;
    if (upper == 0) return;
    // hash set capacity: 2x the candidate count, capped; grown on overflow
    i64 cap = 1024;
    const i64 want = (i64)std::min<unsigned long long>(upper, 1ULL << 29);
    while (cap < 2 * want) cap <<= 1;
    for (;;) {
        R.keys.alloc((size_t)cap);
        R.pos.alloc((size_t)cap);
        if (pool) {
            R.score.alloc((size_t)cap);
            R.plain.alloc((size_t)cap);
        }
        k_set_clear<<<4096, 256, 0, s>>>(R.keys.p, R.pos.p, pool ? R.score.p : nullptr, pool ? R.plain.p : nullptr,
                                         cap);
        PE_LAUNCH_CHECK();
        PE_HIP(hipMemsetAsync(R.cnt.p + 1, 0, sizeof(unsigned long long), s));
        PE_HIP(hipMemsetAsync(R.ovf.p, 0, sizeof(int), s));
        R.S = NewSet{R.keys.p, R.pos.p, (u64)(cap - 1), R.cnt.p + 1, R.ovf.p};
        if (pool) {
            R.S.score = R.score.p;
            R.S.plain = R.plain.p;
        }
        launch_hci<2>(wfn->kind, w, H, dc, eps, eps_plain, R.stride, R.S, nullptr, s, hb.L);
        unsigned long long uniq = 0;
        int of = 0;
        PE_HIP(hipMemcpyAsync(&uniq, R.cnt.p + 1, sizeof(uniq), hipMemcpyDeviceToHost, s));
        PE_HIP(hipMemcpyAsync(&of, R.ovf.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        if (!of && 2 * (i64)uniq <= cap) {
            R.S_unique_host = uniq;
            break;
        }
        PE_REQUIRE(cap < ((i64)1 << 34), PYCI_ENOMEM, "selection hash set too large");
        cap <<= 2;
    }
    // every claimed slot is one new determinant with its first position
    const i64 nnew = (i64)R.S_unique_host;
    R.nnew = nnew;
    if (nnew == 0) return;
    PE_REQUIRE(nnew < ((i64)1 << 31), PYCI_ENOMEM, "too many new determinants in one selection step");
    DevBuf<u64> pos_in((size_t)nnew), slot_in((size_t)nnew), pos_out((size_t)nnew);
    R.slots.alloc((size_t)nnew);
    PE_HIP(hipMemsetAsync(R.cnt.p, 0, sizeof(unsigned long long), s));
    k_set_owners<<<(unsigned)std::min<i64>(4096, grid_for(cap, 256)), 256, 0, s>>>(R.S, cap, R.cnt.p, pos_in.p,
                                                                                   slot_in.p);
    PE_LAUNCH_CHECK();
    // positions are below nd * stride
    int end_bit = 1;
    while (end_bit < 64 && (long double)((unsigned long long)1 << end_bit) < (long double)nd * (long double)R.stride)
        ++end_bit;
    size_t tmp_bytes = 0;
    PE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, pos_in.p, pos_out.p, slot_in.p, R.slots.p,
                                              (int)nnew, 0, end_bit, s));
    DevBuf<unsigned char> tmp(tmp_bytes);
    PE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, pos_in.p, pos_out.p, slot_in.p, R.slots.p,
                                              (int)nnew, 0, end_bit, s));
    PE_HIP(hipStreamSynchronize(s));  // the sort's buffers leave scope
}

i64 add_hci_device(pyci_ham *ham, pyci_wfn *wfn, const double *coeffs_host, double eps) {
    pyci_ctx *ctx = wfn->ctx;
    hipStream_t s = ctx->stream;
    const i64 nd = wfn->ndet;
    if (nd == 0) return 0;
    DevBuf<double> dc((size_t)nd);
    PE_HIP(hipMemcpyAsync(dc.p, coeffs_host, sizeof(double) * nd, hipMemcpyHostToDevice, s));
    HciRun R;
    hci_collect(ham, wfn, dc.p, eps, eps, false, R);
    const i64 nnew = R.nnew;
    if (nnew == 0) return 0;
    // grow the device det array, then write the new determinants in place
    const i64 need = (nd + nnew) * wfn->nkey;
    if ((size_t)need > wfn->dets.n) {
        DevBuf<u64> nb((size_t)(need + need / 2));
        PE_HIP(hipMemcpyAsync(nb.p, wfn->dets.p, sizeof(u64) * nd * wfn->nkey, hipMemcpyDeviceToDevice, s));
        PE_HIP(hipStreamSynchronize(s));
        wfn->dets = std::move(nb);
    }
    k_emit_owners<<<grid_for(nnew, 256), 256, 0, s>>>(view_of(wfn), R.S, R.slots.p, nnew,
                                                      wfn->dets.p + nd * wfn->nkey);
    PE_LAUNCH_CHECK();
    PE_HIP(hipStreamSynchronize(s));
    wfn->ndet = nd + nnew;
    wfn_index_append(wfn, nd);
    PE_HIP(hipStreamSynchronize(s));
    return nnew;
}

// dst[i] = src[i] for i < n, dst[n] = 0 (input of an exclusive scan).
__global__ void k_int_to_i64(const int *src, i64 n, i64 *dst) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) dst[i] = (i < n) ? (i64)src[i] : 0;
}

// Score and plain flag of each pooled candidate, from its slot.
__global__ void k_pool_scores(NewSet S, const u64 *slots, i64 n, double *hb, int *plain) {
    const i64 o = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const u64 slot = slots[o];
    hb[o] = __longlong_as_double((long long)S.score[slot]);
    plain[o] = S.plain[slot];
}

void hci_pool_device(pyci_ham *ham, pyci_wfn *wfn, const double *dc, double eps, double eps_pool, HciPool &P) {
    hipStream_t s = wfn->ctx->stream;
    P.n = 0;
    P.nplain = 0;
    if (wfn->ndet == 0) return;
    HciRun R;
    hci_collect(ham, wfn, dc, eps_pool, eps, true, R);
    if (R.nnew == 0) return;
    P.n = R.nnew;
    P.dets.alloc((size_t)(R.nnew * wfn->nkey));
    P.hb.alloc((size_t)R.nnew);
    P.plain.alloc((size_t)R.nnew);
    const WfnView w = view_of(wfn);
    k_emit_owners<<<grid_for(P.n, 256), 256, 0, s>>>(w, R.S, R.slots.p, P.n, P.dets.p);
    PE_LAUNCH_CHECK();
    k_pool_scores<<<grid_for(P.n, 256), 256, 0, s>>>(R.S, R.slots.p, P.n, P.hb.p, P.plain.p);
    PE_LAUNCH_CHECK();
    // number of candidates plain add_hci(eps) appends
    DevBuf<i64> cnt((size_t)(P.n + 1));
    k_int_to_i64<<<grid_for(P.n + 1, 256), 256, 0, s>>>(P.plain.p, P.n, cnt.p);
    PE_LAUNCH_CHECK();
    P.nplain = exclusive_scan(wfn->ctx, cnt.p, P.n + 1);
    PE_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Epstein-Nesbet PT2 (enpt2.cpp:342-370): the same candidate enumeration and
// screening as add_hci, with every passing external determinant a accumulating
// sum_i H_ai c_i in a device hash set (pass 5), then
// sum_a num_a^2 / (e - H_aa) over the set.
// ---------------------------------------------------------------------------

// H_aa in the reference's order (compute_enpt2_thread_gather, enpt2.cpp:37-70;
// GenCI :237-255): per occupied j, one_mo[j,j], same-spin pairs, then for FullCI
// the opposite-spin pairs; then the beta string's own terms.
__device__ inline double pt2_diag(const HamView &H, u64 up, u64 dn) {
    const int n = H.n;
    const double *one = H.one, *two = H.two;
    double d = 0.0;
    for (u64 x = up; x; x &= x - 1) {
        const int j = __builtin_ctzll(x);
        d += one[j * (n + 1)];
        for (u64 y = x & (x - 1); y; y &= y - 1) {
            const int l = __builtin_ctzll(y);
            d += two[GIX(j, l, j, l)] - two[GIX(j, l, l, j)];
        }
        for (u64 y = dn; y; y &= y - 1) {
            const int l = __builtin_ctzll(y);
            d += two[GIX(j, l, j, l)];
        }
    }
    for (u64 x = dn; x; x &= x - 1) {
        const int j = __builtin_ctzll(x);
        d += one[j * (n + 1)];
        for (u64 y = x & (x - 1); y; y &= y - 1) {
            const int l = __builtin_ctzll(y);
            d += two[GIX(j, l, j, l)] - two[GIX(j, l, l, j)];
        }
    }
    return d;
}

constexpr int PT2_BLOCKS = 1024;

// partial[b] = block b's sum of num^2 / (e - H_aa) over its slots (grid-stride,
// then a fixed-order tree in LDS, so the total does not depend on timing).
__global__ void __launch_bounds__(256) k_pt2_terms(NewSet S, i64 cap, HamView H, double e, double *partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (i64 h = (i64)blockIdx.x * 256 + threadIdx.x; h < cap; h += (i64)gridDim.x * 256) {
        if (S.keys[h] == EMPTY_KEY) continue;
        const double x = S.num[h];
        acc += x * x / (e - pt2_diag(H, S.kdet[2 * h], S.kdet[2 * h + 1]));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_pt2_clear(double *num, u64 *kdet, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        num[i] = 0.0;
        kdet[2 * i] = 0ULL;
        kdet[2 * i + 1] = 0ULL;
    }
}

// (d) -> (d, d): the FullCI set of a DOCI set, in the same order (FullCIWfn(const
// DOCIWfn &), fullciwfn.cpp:26-35).
__global__ void k_doci_to_fullci(const u64 *src, i64 n, u64 *dst) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[2 * i] = dst[2 * i + 1] = src[i];
}

double enpt2_device(pyci_ham *ham, pyci_wfn *wfn_in, const double *coeffs_host, double e, double eps, i64 *nterms) {
    pyci_ctx *ctx = wfn_in->ctx;
    hipStream_t s = ctx->stream;
    PE_REQUIRE(ham->nbasis == wfn_in->nbasis, PYCI_EINVAL, "operator and wave function have different nbasis");
    if (nterms) *nterms = 0;
    const i64 nd = wfn_in->ndet;
    if (nd == 0) return 0.0;
    // a DOCI set enters as the FullCI set of its (d, d) determinants (enpt2.cpp:382-386)
    pyci_wfn tmp;
    pyci_wfn *wfn = wfn_in;
    if (wfn_in->kind == PYCI_WFN_DOCI) {
        tmp.ctx = ctx;
        tmp.kind = PYCI_WFN_FULLCI;
        tmp.nbasis = wfn_in->nbasis;
        tmp.nup = tmp.ndn = wfn_in->nup;
        tmp.nkey = 2;
        tmp.nsd = binom_host(tmp.nbasis, tmp.ndn);
        tmp.ndet = nd;
        tmp.dets.alloc((size_t)(2 * nd));
        k_doci_to_fullci<<<grid_for(nd, 256), 256, 0, s>>>(wfn_in->dets.p, nd, tmp.dets.p);
        PE_LAUNCH_CHECK();
        wfn_build_index(&tmp);
        wfn = &tmp;
    }
    if (wfn->kind == PYCI_WFN_FULLCI) {
        long double tot = (long double)binom_host(wfn->nbasis, wfn->nup) * (long double)binom_host(wfn->nbasis, wfn->ndn);
        PE_REQUIRE(tot < 9.0e18L, PYCI_EUNSUPPORTED, "determinant space too large for the GPU selection keys");
    }
    DevBuf<double> dc((size_t)nd);
    PE_HIP(hipMemcpyAsync(dc.p, coeffs_host, sizeof(double) * nd, hipMemcpyHostToDevice, s));
    const WfnView w = view_of(wfn);
    const HamView H = view_of(ham);
    DevBuf<unsigned long long> cnt(2);
    DevBuf<int> ovf(1);
    PE_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * 2, s));
    NewSet S{nullptr, nullptr, 0, cnt.p + 1, ovf.p};
    // pass 1: passing external candidates (with repeats) bound the set size
    HbTables hb;
    if (wfn->kind == PYCI_WFN_GENCI) hb.build(H, s);
    launch_hci<1>(wfn->kind, w, H, dc.p, eps, eps, 0, S, cnt.p, s, hb.L);
    unsigned long long upper = 0;
    PE_HIP(hipMemcpyAsync(&upper, cnt.p, sizeof(upper), hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    if (upper == 0) return 0.0;
    i64 cap = 1024;
    const i64 want = (i64)std::min<unsigned long long>(upper, 1ULL << 29);
    while (cap < 2 * want) cap <<= 1;
    DevBuf<u64> keys, pos, kdet;
    DevBuf<double> num;
    unsigned long long uniq = 0;
    for (;;) {
        keys.alloc((size_t)cap);
        pos.alloc((size_t)cap);
        num.alloc((size_t)cap);
        kdet.alloc((size_t)(2 * cap));
        k_set_clear<<<4096, 256, 0, s>>>(keys.p, pos.p, nullptr, nullptr, cap);
        PE_LAUNCH_CHECK();
        k_pt2_clear<<<4096, 256, 0, s>>>(num.p, kdet.p, cap);
        PE_LAUNCH_CHECK();
        PE_HIP(hipMemsetAsync(cnt.p + 1, 0, sizeof(unsigned long long), s));
        PE_HIP(hipMemsetAsync(ovf.p, 0, sizeof(int), s));
        S = NewSet{keys.p, pos.p, (u64)(cap - 1), cnt.p + 1, ovf.p};
        S.num = num.p;
        S.kdet = kdet.p;
        launch_hci<5>(wfn->kind, w, H, dc.p, eps, eps, 0, S, nullptr, s, hb.L);
        int of = 0;
        PE_HIP(hipMemcpyAsync(&uniq, cnt.p + 1, sizeof(uniq), hipMemcpyDeviceToHost, s));
        PE_HIP(hipMemcpyAsync(&of, ovf.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        if (!of && 2 * (i64)uniq <= cap) break;
        PE_REQUIRE(cap < ((i64)1 << 34), PYCI_ENOMEM, "ENPT2 hash set too large");
        cap <<= 2;
    }
    DevBuf<double> part(PT2_BLOCKS);
    k_pt2_terms<<<PT2_BLOCKS, 256, 0, s>>>(S, cap, H, e, part.p);
    PE_LAUNCH_CHECK();
    std::vector<double> hp(PT2_BLOCKS);
    PE_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * PT2_BLOCKS, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    double corr = 0.0;
    for (double v : hp) corr += v;
    if (nterms) *nterms = (i64)uniq;
    return corr;
}

}  // namespace pe
