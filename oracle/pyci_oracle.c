/*
 * CPU ORACLE for the configuration-interaction hot path.
 *
 * This file is TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline. The product (pyci_amd) never links or calls it.
 *
 * It is a plain-C restatement of the reference algorithms (PyCI v0.6.1), one
 * determinant word per spin (nbasis <= 64 spatial orbitals for DOCI/FullCI,
 * <= 64 spin-orbitals for GenCI). Each function cites the reference lines
 * it follows. Pins for every function are listed in oracle/README.md and
 * exercised by tests/test_oracle_pins.py.
 *
 * Determinant layout: DOCI and GenCI use one uint64 per determinant; FullCI uses
 * two (alpha word, then beta word), matching the reference's `dets` vectors
 * (onespinwfn.cpp:57-63, twospinwfn.cpp:57-63).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KIND_DOCI 0
#define KIND_FULLCI 1
#define KIND_GENCI 2

typedef uint64_t u64;
typedef int64_t i64;

/* ------------------------------------------------------------------------ */
/* Bit primitives (common.cpp:41-66, 85-113, 115-153, 155-261)               */
/* ------------------------------------------------------------------------ */

/* Exact binomial coefficient for n <= 66 (common.cpp:41-66 gives the same
 * values in this range). */
i64 oracle_binomial(i64 n, i64 k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 b = 1;
    for (i64 d = 1; d <= k; ++d) b = b * (unsigned __int128)(n - k + d) / (unsigned __int128)d;
    return (i64)b;
}

/* fill_occs / fill_virs for one word (common.cpp:85-113): ascending bit order. */
static int occs_of(u64 det, i64 *occs) {
    int n = 0;
    while (det) {
        int p = __builtin_ctzll(det);
        occs[n++] = p;
        det &= det - 1;
    }
    return n;
}

static int virs_of(u64 det, i64 nbasis, i64 *virs) {
    u64 mask = (nbasis >= 64) ? ~0ULL : ((1ULL << nbasis) - 1ULL);
    return occs_of(det ^ mask, virs);
}

/* Number of set bits strictly between positions lo and hi (lo < hi). */
static int between(u64 det, i64 lo, i64 hi) {
    u64 below_hi = (hi >= 64) ? ~0ULL : ((1ULL << hi) - 1ULL);
    u64 above_lo = (lo >= 63) ? 0ULL : ~((1ULL << (lo + 1)) - 1ULL);
    return __builtin_popcountll(det & below_hi & above_lo);
}

/* phase_single_det (common.cpp:163-194): (-1)^(occupied strictly between i, a). */
i64 oracle_phase_single(u64 det, i64 i, i64 a) {
    i64 lo = i < a ? i : a, hi = i < a ? a : i;
    return (between(det, lo, hi) & 1) ? -1 : 1;
}

/* phase_double_det (common.cpp:196-261): both single counts on the ORIGINAL
 * determinant, plus one if (i2 < a1 || i1 > a2). */
i64 oracle_phase_double(u64 det, i64 i1, i64 i2, i64 a1, i64 a2) {
    int n = 0;
    n += between(det, i1 < a1 ? i1 : a1, i1 < a1 ? a1 : i1);
    n += between(det, i2 < a2 ? i2 : a2, i2 < a2 ? a2 : i2);
    if (i2 < a1 || i1 > a2) ++n;
    return (n & 1) ? -1 : 1;
}

/* Colex rank (common.cpp:125-139): sum over the k-th set bit at position p of C(p, k+1). */
i64 oracle_rank_colex(i64 nbasis, i64 nocc, u64 det) {
    (void)nbasis;
    i64 r = 0, k = 0;
    while (det && k < nocc) {
        int p = __builtin_ctzll(det);
        ++k;
        r += oracle_binomial(p, k);
        det &= det - 1;
    }
    return r;
}

/* Colex unrank (common.cpp:141-153). */
u64 oracle_unrank_colex(i64 nbasis, i64 nocc, i64 rank) {
    u64 det = 0;
    i64 n = nbasis;
    for (i64 k = nocc; k >= 1; --k) {
        /* largest p < n with C(p, k) <= rank */
        i64 p = n - 1;
        while (oracle_binomial(p, k) > rank) --p;
        det |= 1ULL << p;
        rank -= oracle_binomial(p, k);
        n = p;
    }
    return det;
}

/* ------------------------------------------------------------------------ */
/* Insertion-ordered determinant set (stands for Wfn::dets + Wfn::dict)      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nkey;     /* words per determinant: 1 or 2 */
    i64 n, cap;   /* number of dets, capacity */
    u64 *keys;    /* n * nkey */
    i64 hcap;     /* power of two */
    i64 *slots;   /* -1 = empty, else det index */
} dset;

static u64 mix64(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static u64 hkey(const u64 *k, int nkey) {
    u64 h = mix64(k[0] + 0x9e3779b97f4a7c15ULL);
    if (nkey == 2) h = mix64(h ^ (k[1] + 0x632be59bd9b4e019ULL));
    return h;
}

static void dset_init(dset *s, int nkey, i64 cap) {
    s->nkey = nkey;
    s->n = 0;
    s->cap = cap > 16 ? cap : 16;
    s->keys = (u64 *)malloc(sizeof(u64) * s->cap * nkey);
    s->hcap = 32;
    while (s->hcap < 2 * s->cap) s->hcap <<= 1;
    s->slots = (i64 *)malloc(sizeof(i64) * s->hcap);
    for (i64 i = 0; i < s->hcap; ++i) s->slots[i] = -1;
}

static void dset_free(dset *s) {
    free(s->keys);
    free(s->slots);
}

static i64 dset_find(const dset *s, const u64 *k) {
    i64 m = s->hcap - 1, h = (i64)(hkey(k, s->nkey) & (u64)m);
    for (;;) {
        i64 v = s->slots[h];
        if (v < 0) return -1;
        const u64 *kk = s->keys + v * s->nkey;
        if (kk[0] == k[0] && (s->nkey == 1 || kk[1] == k[1])) return v;
        h = (h + 1) & m;
    }
}

static void dset_rehash(dset *s) {
    for (i64 i = 0; i < s->hcap; ++i) s->slots[i] = -1;
    i64 m = s->hcap - 1;
    for (i64 v = 0; v < s->n; ++v) {
        i64 h = (i64)(hkey(s->keys + v * s->nkey, s->nkey) & (u64)m);
        while (s->slots[h] >= 0) h = (h + 1) & m;
        s->slots[h] = v;
    }
}

/* add_det (onespinwfn.cpp:136-143, twospinwfn.cpp:140-147): returns the new index or -1. */
static i64 dset_add(dset *s, const u64 *k) {
    if (dset_find(s, k) >= 0) return -1;
    if (s->n == s->cap) {
        s->cap *= 2;
        s->keys = (u64 *)realloc(s->keys, sizeof(u64) * s->cap * s->nkey);
    }
    if (2 * (s->n + 1) > s->hcap) {
        s->hcap *= 2;
        s->slots = (i64 *)realloc(s->slots, sizeof(i64) * s->hcap);
        dset_rehash(s);
    }
    memcpy(s->keys + s->n * s->nkey, k, sizeof(u64) * s->nkey);
    i64 m = s->hcap - 1, h = (i64)(hkey(k, s->nkey) & (u64)m);
    while (s->slots[h] >= 0) h = (h + 1) & m;
    s->slots[h] = s->n;
    return s->n++;
}

static void dset_from(dset *s, int nkey, i64 ndet, const u64 *dets) {
    dset_init(s, nkey, ndet + 16);
    for (i64 i = 0; i < ndet; ++i) dset_add(s, dets + i * nkey);
}

/* ------------------------------------------------------------------------ */
/* add_all_dets (onespinwfn.cpp:173-217, twospinwfn.cpp:181-245)             */
/* ------------------------------------------------------------------------ */

/* Full space in colex order: one-spin det i = unrank(i); two-spin det
 * i_up * C(nb, ndn) + i_dn = (unrank_up(i_up), unrank_dn(i_dn)). */
i64 oracle_all_dets(int kind, i64 nbasis, i64 nup, i64 ndn, u64 *out) {
    if (kind == KIND_FULLCI) {
        i64 mu = oracle_binomial(nbasis, nup), md = oracle_binomial(nbasis, ndn);
        for (i64 iu = 0; iu < mu; ++iu) {
            u64 up = oracle_unrank_colex(nbasis, nup, iu);
            for (i64 id = 0; id < md; ++id) {
                out[2 * (iu * md + id)] = up;
                out[2 * (iu * md + id) + 1] = oracle_unrank_colex(nbasis, ndn, id);
            }
        }
        return mu * md;
    }
    i64 m = oracle_binomial(nbasis, nup);
    for (i64 i = 0; i < m; ++i) out[i] = oracle_unrank_colex(nbasis, nup, i);
    return m;
}

/* ------------------------------------------------------------------------ */
/* add_excited_dets (onespinwfn.cpp:219-244, twospinwfn.cpp:247-265,325-342)  */
/* ------------------------------------------------------------------------ */

/* All rank-e excitations of one spin string, in the reference order:
 * virtual combinations (outer, colex) x occupied combinations (inner, colex). */
static i64 one_spin_excitations(u64 ref, i64 nbasis, i64 e, u64 *out) {
    i64 occs[64], virs[64];
    int no = occs_of(ref, occs), nv = virs_of(ref, nbasis, virs);
    i64 nvc = oracle_binomial(nv, e), noc = oracle_binomial(no, e), cnt = 0;
    for (i64 iv = 0; iv < nvc; ++iv) {
        u64 vsel = oracle_unrank_colex(nv, e, iv);
        for (i64 io = 0; io < noc; ++io) {
            u64 osel = oracle_unrank_colex(no, e, io);
            u64 det = ref;
            u64 vs = vsel, os = osel;
            for (i64 k = 0; k < e; ++k) {
                int pv = __builtin_ctzll(vs), po = __builtin_ctzll(os);
                vs &= vs - 1;
                os &= os - 1;
                det &= ~(1ULL << occs[po]);
                det |= 1ULL << virs[pv];
            }
            out[cnt++] = det;
        }
    }
    return cnt;
}

/* Appends to (ndet, dets) all rank-exc excitations of ref not already present.
 * Returns the new ndet; out must hold the result (capacity cap dets). */
i64 oracle_add_excited_dets(int kind, i64 nbasis, i64 nup, i64 ndn, i64 ndet, const u64 *dets,
                            const u64 *ref, i64 exc, u64 *out, i64 cap) {
    int nkey = (kind == KIND_FULLCI) ? 2 : 1;
    dset s;
    dset_from(&s, nkey, ndet, dets);
    if (kind != KIND_FULLCI) {
        u64 *buf = (u64 *)malloc(sizeof(u64) * (oracle_binomial(nup, exc) * oracle_binomial(nbasis - nup, exc) + 1));
        i64 n = one_spin_excitations(ref[0], nbasis, exc, buf);
        for (i64 i = 0; i < n; ++i) dset_add(&s, buf + i);
        free(buf);
    } else if (exc == 0) {
        dset_add(&s, ref);
    } else {
        i64 maxup = nup < nbasis - nup ? nup : nbasis - nup;
        i64 maxdn = ndn < nbasis - ndn ? ndn : nbasis - ndn;
        i64 a = exc < maxup ? exc : maxup, b = exc - a;
        while (a >= 0 && b <= maxdn) {
            i64 nu_ = oracle_binomial(nup, a) * oracle_binomial(nbasis - nup, a);
            i64 nd_ = oracle_binomial(ndn, b) * oracle_binomial(nbasis - ndn, b);
            u64 *bu = (u64 *)malloc(sizeof(u64) * (nu_ + 1)), *bd = (u64 *)malloc(sizeof(u64) * (nd_ + 1));
            one_spin_excitations(ref[0], nbasis, a, bu);
            one_spin_excitations(ref[1], nbasis, b, bd);
            for (i64 i = 0; i < nu_; ++i)
                for (i64 j = 0; j < nd_; ++j) {
                    u64 k[2] = {bu[i], bd[j]};
                    dset_add(&s, k);
                }
            free(bu);
            free(bd);
            --a;
            ++b;
        }
    }
    i64 n = s.n < cap ? s.n : cap;
    memcpy(out, s.keys, sizeof(u64) * n * nkey);
    i64 res = s.n;
    dset_free(&s);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Sparse CI operator (sparseop.cpp:186-502)                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    i64 row0;               /* global index of the first stored row */
    i64 nrow, ncol, nnz, cap; /* nrow = number of stored rows */
    double *data;
    i64 *indices;
    i64 *indptr;
} oracle_csr;

static void csr_push(oracle_csr *c, i64 j, double v) {
    if (c->nnz == c->cap) {
        c->cap = c->cap * 3 / 2 + 1024;
        c->data = (double *)realloc(c->data, sizeof(double) * c->cap);
        c->indices = (i64 *)realloc(c->indices, sizeof(i64) * c->cap);
    }
    c->data[c->nnz] = v;
    c->indices[c->nnz++] = j;
}

/* sort_row (sparseop.cpp:214-218): ascending column; columns are unique. */
static void sort_row(oracle_csr *c, i64 start, i64 end) {
    for (i64 a = start + 1; a < end; ++a) {
        i64 j = c->indices[a];
        double v = c->data[a];
        i64 b = a - 1;
        while (b >= start && c->indices[b] > j) {
            c->indices[b + 1] = c->indices[b];
            c->data[b + 1] = c->data[b];
            --b;
        }
        c->indices[b + 1] = j;
        c->data[b + 1] = v;
    }
}

typedef struct {
    int kind;
    i64 n1, nup, ndn;
    const double *one_mo, *two_mo;
    double *h, *v, *w; /* seniority-zero integrals (squantop.cpp:174-182) */
} hamv;

static void ham_init(hamv *H, int kind, i64 nbasis, i64 nup, i64 ndn, const double *one_mo, const double *two_mo) {
    H->kind = kind;
    H->n1 = nbasis;
    H->nup = nup;
    H->ndn = ndn;
    H->one_mo = one_mo;
    H->two_mo = two_mo;
    i64 n = nbasis, n2 = n * n, n3 = n2 * n;
    H->h = (double *)malloc(sizeof(double) * n);
    H->v = (double *)malloc(sizeof(double) * n2);
    H->w = (double *)malloc(sizeof(double) * n2);
    for (i64 i = 0; i < n; ++i) {
        H->h[i] = one_mo[i * (n + 1)];
        for (i64 j = 0; j < n; ++j) {
            H->v[i * n + j] = two_mo[i * n3 + i * n2 + j * n + j];
            H->w[i * n + j] = two_mo[i * n3 + j * n2 + i * n + j] * 2 - two_mo[i * n3 + j * n2 + j * n + i];
        }
    }
}

static void ham_free(hamv *H) {
    free(H->h);
    free(H->v);
    free(H->w);
}

#define G(p, q, r, s) (H->two_mo[(p) * n3 + (q) * n2 + (r) * n1 + (s)])

/* add_row(DOCIWfn) (sparseop.cpp:220-258). */
static void row_doci(oracle_csr *c, const hamv *H, const dset *s, i64 idet, i64 jmin) {
    i64 n = H->n1, occs[64], virs[64];
    u64 det = s->keys[idet];
    int no = occs_of(det, occs), nv = virs_of(det, n, virs);
    double val1 = 0.0, val2 = 0.0;
    for (int i = 0; i < no; ++i) {
        i64 k = occs[i];
        val1 += H->v[k * (n + 1)];
        val2 += H->h[k];
        for (int j = i + 1; j < no; ++j) val2 += H->w[k * n + occs[j]];
        for (int j = 0; j < nv; ++j) {
            i64 l = virs[j];
            u64 e = (det & ~(1ULL << k)) | (1ULL << l);
            i64 jdet = dset_find(s, &e);
            if (jdet != -1 && jdet < jmin && jdet < c->ncol) csr_push(c, jdet, H->v[k * n + l]);
        }
    }
    if (idet < c->ncol) csr_push(c, idet, val1 + val2 * 2);
}

/* add_row(FullCIWfn) (sparseop.cpp:260-427). */
static void row_fullci(oracle_csr *c, const hamv *H, const dset *s, i64 idet, i64 jmin) {
    i64 n1 = H->n1, n2 = n1 * n1, n3 = n2 * n1;
    i64 ou[64], od[64], vu[64], vd[64];
    u64 up = s->keys[2 * idet], dn = s->keys[2 * idet + 1];
    int nou = occs_of(up, ou), nod = occs_of(dn, od);
    int nvu = virs_of(up, n1, vu), nvd = virs_of(dn, n1, vd);
    double val1, val2 = 0.0;
    for (int i = 0; i < nou; ++i) {
        i64 ii = ou[i];
        val2 += H->one_mo[(n1 + 1) * ii];
        for (int k = i + 1; k < nou; ++k) {
            i64 kk = ou[k];
            val2 += G(ii, kk, ii, kk) - G(ii, kk, kk, ii);
        }
        for (int k = 0; k < nod; ++k) {
            i64 kk = od[k];
            val2 += G(ii, kk, ii, kk);
        }
        for (int j = 0; j < nvu; ++j) {
            i64 jj = vu[j];
            u64 eu = (up & ~(1ULL << ii)) | (1ULL << jj);
            i64 sign_up = oracle_phase_single(up, ii, jj);
            u64 key[2] = {eu, dn};
            i64 jdet = dset_find(s, key);
            if (jdet != -1 && jdet < jmin && jdet < c->ncol) {
                val1 = H->one_mo[n1 * ii + jj];
                for (int k = 0; k < nou; ++k) {
                    i64 kk = ou[k];
                    val1 += G(ii, kk, jj, kk) - G(ii, kk, kk, jj);
                }
                for (int k = 0; k < nod; ++k) {
                    i64 kk = od[k];
                    val1 += G(ii, kk, jj, kk);
                }
                csr_push(c, jdet, sign_up * val1);
            }
            for (int k = 0; k < nod; ++k) {
                i64 kk = od[k];
                for (int l = 0; l < nvd; ++l) {
                    i64 ll = vd[l];
                    u64 k2[2] = {eu, (dn & ~(1ULL << kk)) | (1ULL << ll)};
                    jdet = dset_find(s, k2);
                    if (jdet != -1 && jdet < jmin && jdet < c->ncol)
                        csr_push(c, jdet, (double)(sign_up * oracle_phase_single(dn, kk, ll)) * G(ii, kk, jj, ll));
                }
            }
            for (int k = i + 1; k < nou; ++k) {
                i64 kk = ou[k];
                for (int l = j + 1; l < nvu; ++l) {
                    i64 ll = vu[l];
                    u64 k2[2] = {(eu & ~(1ULL << kk)) | (1ULL << ll), dn};
                    jdet = dset_find(s, k2);
                    if (jdet != -1 && jdet < jmin && jdet < c->ncol)
                        csr_push(c, jdet, (double)oracle_phase_double(up, ii, kk, jj, ll) * (G(ii, kk, jj, ll) - G(ii, kk, ll, jj)));
                }
            }
        }
    }
    for (int i = 0; i < nod; ++i) {
        i64 ii = od[i];
        val2 += H->one_mo[(n1 + 1) * ii];
        for (int k = i + 1; k < nod; ++k) {
            i64 kk = od[k];
            val2 += G(ii, kk, ii, kk) - G(ii, kk, kk, ii);
        }
        for (int j = 0; j < nvd; ++j) {
            i64 jj = vd[j];
            u64 ed = (dn & ~(1ULL << ii)) | (1ULL << jj);
            u64 key[2] = {up, ed};
            i64 jdet = dset_find(s, key);
            if (jdet != -1 && jdet < jmin && jdet < c->ncol) {
                val1 = H->one_mo[n1 * ii + jj];
                for (int k = 0; k < nou; ++k) {
                    i64 kk = ou[k];
                    val1 += G(ii, kk, jj, kk);
                }
                for (int k = 0; k < nod; ++k) {
                    i64 kk = od[k];
                    val1 += G(ii, kk, jj, kk) - G(ii, kk, kk, jj);
                }
                csr_push(c, jdet, oracle_phase_single(dn, ii, jj) * val1);
            }
            for (int k = i + 1; k < nod; ++k) {
                i64 kk = od[k];
                for (int l = j + 1; l < nvd; ++l) {
                    i64 ll = vd[l];
                    u64 k2[2] = {up, (ed & ~(1ULL << kk)) | (1ULL << ll)};
                    jdet = dset_find(s, k2);
                    if (jdet != -1 && jdet < jmin && jdet < c->ncol)
                        csr_push(c, jdet, (double)oracle_phase_double(dn, ii, kk, jj, ll) * (G(ii, kk, jj, ll) - G(ii, kk, ll, jj)));
                }
            }
        }
    }
    if (idet < c->ncol) csr_push(c, idet, val2);
}

/* add_row(GenCIWfn) (sparseop.cpp:429-502). */
static void row_genci(oracle_csr *c, const hamv *H, const dset *s, i64 idet, i64 jmin) {
    i64 n1 = H->n1, n2 = n1 * n1, n3 = n2 * n1;
    i64 occs[64], virs[64];
    u64 det = s->keys[idet];
    int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
    double val1, val2 = 0.0;
    for (int i = 0; i < no; ++i) {
        i64 ii = occs[i];
        val2 += H->one_mo[(n1 + 1) * ii];
        for (int k = i + 1; k < no; ++k) {
            i64 kk = occs[k];
            val2 += G(ii, kk, ii, kk) - G(ii, kk, kk, ii);
        }
        for (int j = 0; j < nv; ++j) {
            i64 jj = virs[j];
            u64 e = (det & ~(1ULL << ii)) | (1ULL << jj);
            i64 jdet = dset_find(s, &e);
            if (jdet != -1 && jdet < jmin && jdet < c->ncol) {
                val1 = H->one_mo[n1 * ii + jj];
                for (int k = 0; k < no; ++k) {
                    i64 kk = occs[k];
                    val1 += G(ii, kk, jj, kk) - G(ii, kk, kk, jj);
                }
                csr_push(c, jdet, oracle_phase_single(det, ii, jj) * val1);
            }
            for (int k = i + 1; k < no; ++k) {
                i64 kk = occs[k];
                for (int l = j + 1; l < nv; ++l) {
                    i64 ll = virs[l];
                    u64 e2 = (e & ~(1ULL << kk)) | (1ULL << ll);
                    jdet = dset_find(s, &e2);
                    if (jdet != -1 && jdet < jmin && jdet < c->ncol)
                        csr_push(c, jdet, (double)oracle_phase_double(det, ii, kk, jj, ll) * (G(ii, kk, jj, ll) - G(ii, kk, ll, jj)));
                }
            }
        }
    }
    if (idet < c->ncol) csr_push(c, idet, val2);
}

/* SparseOp(ham, wfn, nrow, ncol, symmetric) -> update(..., startrow = row0)
 * (sparseop.cpp:57-63, 186-201). With symmetric != 0 only columns < row are
 * kept (jmin = idet), plus the diagonal. Rows row0..nrow-1 are built. */
oracle_csr *oracle_csr_build(int kind, i64 nbasis, i64 nup, i64 ndn, i64 ndet, const u64 *dets,
                             const double *one_mo, const double *two_mo, i64 row0, i64 nrow, i64 ncol,
                             int symmetric) {
    int nkey = (kind == KIND_FULLCI) ? 2 : 1;
    dset s;
    dset_from(&s, nkey, ndet, dets);
    hamv H;
    ham_init(&H, kind, nbasis, nup, ndn, one_mo, two_mo);
    oracle_csr *c = (oracle_csr *)calloc(1, sizeof(oracle_csr));
    c->row0 = row0;
    c->nrow = nrow - row0;
    c->ncol = ncol;
    c->indptr = (i64 *)malloc(sizeof(i64) * (nrow - row0 + 1));
    c->indptr[0] = 0;
    for (i64 idet = row0; idet < nrow; ++idet) {
        i64 jmin = symmetric ? idet : INT64_MAX;
        i64 start = c->nnz;
        if (kind == KIND_DOCI) row_doci(c, &H, &s, idet, jmin);
        else if (kind == KIND_FULLCI) row_fullci(c, &H, &s, idet, jmin);
        else row_genci(c, &H, &s, idet, jmin);
        sort_row(c, start, c->nnz);
        c->indptr[idet - row0 + 1] = c->nnz;
    }
    ham_free(&H);
    dset_free(&s);
    return c;
}

i64 oracle_csr_nnz(const oracle_csr *c) { return c->nnz; }

void oracle_csr_get(const oracle_csr *c, double *data, i64 *indices, i64 *indptr) {
    memcpy(data, c->data, sizeof(double) * c->nnz);
    memcpy(indices, c->indices, sizeof(i64) * c->nnz);
    memcpy(indptr, c->indptr, sizeof(i64) * (c->nrow + 1));
}

void oracle_csr_free(oracle_csr *c) {
    if (!c) return;
    free(c->data);
    free(c->indices);
    free(c->indptr);
    free(c);
}

/* perform_op / perform_op_symm (sparseop.cpp:96-112): y = A x, or with a
 * lower-stored symmetric matrix y = (L + L^T - D) x. ecore is not added. */
/* Rows are global rows row0 .. row0+nrow-1; x has ncol entries and y must hold
 * max(row0+nrow, ncol) entries (the transposed half writes y[col]). */
void oracle_csr_matvec(const oracle_csr *c, int symmetric, const double *x, double *y) {
    for (i64 i = 0; i < c->nrow; ++i) y[c->row0 + i] = 0.0;
    for (i64 i = 0; i < c->nrow; ++i) {
        const i64 gi = c->row0 + i;
        double acc = 0.0;
        for (i64 p = c->indptr[i]; p < c->indptr[i + 1]; ++p) {
            i64 j = c->indices[p];
            acc += c->data[p] * x[j];
            if (symmetric && j != gi) y[j] += c->data[p] * x[gi];
        }
        y[gi] += acc;
    }
}

/* ------------------------------------------------------------------------ */
/* Heat-bath selection (hci.cpp:49-238, driver 252-279)                      */
/* ------------------------------------------------------------------------ */

/* Serial restatement: determinants are appended in FIRST-OCCURRENCE order of
 * the single-threaded enumeration (source det ascending, reference loop order).
 * The reference's own order follows phmap iteration and is not reproducible;
 * the set of added determinants is the same. Returns the number added; the new
 * determinants are written to out (capacity cap). */
i64 oracle_add_hci(int kind, i64 nbasis, i64 nup, i64 ndn, i64 ndet, const u64 *dets, const double *one_mo,
                   const double *two_mo, const double *coeffs, double eps, u64 *out, i64 cap) {
    int nkey = (kind == KIND_FULLCI) ? 2 : 1;
    dset s, t;
    dset_from(&s, nkey, ndet, dets);
    dset_init(&t, nkey, 1024);
    hamv H;
    ham_init(&H, kind, nbasis, nup, ndn, one_mo, two_mo);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1;
    i64 occs[64], virs[64], ou[64], od[64], vu[64], vd[64];
#define TRY_ADD(K)                                           \
    do {                                                     \
        if (dset_find(&s, (K)) == -1) dset_add(&t, (K));     \
    } while (0)
    for (i64 idet = 0; idet < ndet; ++idet) {
        double eps_i = eps / fabs(coeffs[idet]);
        if (kind == KIND_DOCI) {
            u64 det = dets[idet];
            int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
            for (int i = 0; i < no; ++i)
                for (int j = 0; j < nv; ++j) {
                    i64 k = occs[i], l = virs[j];
                    if (fabs(H.v[k * n1 + l]) > eps_i) {
                        u64 e = (det & ~(1ULL << k)) | (1ULL << l);
                        TRY_ADD(&e);
                    }
                }
        } else if (kind == KIND_GENCI) {
            const hamv *HH = &H;
#define GG(p, q, r, s_) (HH->two_mo[(p) * n3 + (q) * n2 + (r) * n1 + (s_)])
            u64 det = dets[idet];
            int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
            for (int i = 0; i < no; ++i) {
                i64 ii = occs[i];
                for (int j = 0; j < nv; ++j) {
                    i64 jj = virs[j];
                    u64 e = (det & ~(1ULL << ii)) | (1ULL << jj);
                    double val = one_mo[n1 * ii + jj];
                    for (int k = 0; k < no; ++k) val += GG(ii, occs[k], jj, occs[k]) - GG(ii, occs[k], occs[k], jj);
                    if (fabs(val) > eps_i) TRY_ADD(&e);
                    for (int k = i + 1; k < no; ++k)
                        for (int l = j + 1; l < nv; ++l) {
                            i64 kk = occs[k], ll = virs[l];
                            val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                            if (fabs(val) > eps_i) {
                                u64 e2 = (e & ~(1ULL << kk)) | (1ULL << ll);
                                TRY_ADD(&e2);
                            }
                        }
                }
            }
        } else {
            const hamv *HH = &H;
            u64 up = dets[2 * idet], dn = dets[2 * idet + 1];
            int nou = occs_of(up, ou), nod = occs_of(dn, od), nvu = virs_of(up, n1, vu), nvd = virs_of(dn, n1, vd);
            for (int i = 0; i < nou; ++i) {
                i64 ii = ou[i];
                for (int j = 0; j < nvu; ++j) {
                    i64 jj = vu[j];
                    u64 eu = (up & ~(1ULL << ii)) | (1ULL << jj);
                    double val = one_mo[n1 * ii + jj];
                    for (int k = 0; k < nou; ++k) val += GG(ii, ou[k], jj, ou[k]) - GG(ii, ou[k], ou[k], jj);
                    for (int k = 0; k < nod; ++k) val += GG(ii, od[k], jj, od[k]);
                    if (fabs(val) > eps_i) {
                        u64 key[2] = {eu, dn};
                        TRY_ADD(key);
                    }
                    for (int k = 0; k < nod; ++k)
                        for (int l = 0; l < nvd; ++l) {
                            i64 kk = od[k], ll = vd[l];
                            val = GG(ii, kk, jj, ll);
                            if (fabs(val) > eps_i) {
                                u64 key[2] = {eu, (dn & ~(1ULL << kk)) | (1ULL << ll)};
                                TRY_ADD(key);
                            }
                        }
                    for (int k = i + 1; k < nou; ++k)
                        for (int l = j + 1; l < nvu; ++l) {
                            i64 kk = ou[k], ll = vu[l];
                            val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                            if (fabs(val) > eps_i) {
                                u64 key[2] = {(eu & ~(1ULL << kk)) | (1ULL << ll), dn};
                                TRY_ADD(key);
                            }
                        }
                }
            }
            for (int i = 0; i < nod; ++i) {
                i64 ii = od[i];
                for (int j = 0; j < nvd; ++j) {
                    i64 jj = vd[j];
                    u64 ed = (dn & ~(1ULL << ii)) | (1ULL << jj);
                    double val = one_mo[n1 * ii + jj];
                    for (int k = 0; k < nou; ++k) val += GG(ii, ou[k], jj, ou[k]);
                    for (int k = 0; k < nod; ++k) val += GG(ii, od[k], jj, od[k]) - GG(ii, od[k], od[k], jj);
                    if (fabs(val) > eps_i) {
                        u64 key[2] = {up, ed};
                        TRY_ADD(key);
                    }
                    for (int k = i + 1; k < nod; ++k)
                        for (int l = j + 1; l < nvd; ++l) {
                            i64 kk = od[k], ll = vd[l];
                            val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                            if (fabs(val) > eps_i) {
                                u64 key[2] = {up, (ed & ~(1ULL << kk)) | (1ULL << ll)};
                                TRY_ADD(key);
                            }
                        }
                }
            }
#undef GG
        }
    }
#undef TRY_ADD
    i64 n = t.n < cap ? t.n : cap;
    memcpy(out, t.keys, sizeof(u64) * n * nkey);
    i64 res = t.n;
    ham_free(&H);
    dset_free(&s);
    dset_free(&t);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Reduced density matrices (rdm.cpp:20-65, 269-530; GenCI by correct math)  */
/* ------------------------------------------------------------------------ */

/* compute_rdms(DOCIWfn) (rdm.cpp:20-65): d0 (n,n), d2 (n,n). */
void oracle_rdms_doci(i64 nbasis, i64 ndet, const u64 *dets, const double *coeffs, double *d0, double *d2) {
    dset s;
    dset_from(&s, 1, ndet, dets);
    i64 n = nbasis, occs[64], virs[64];
    memset(d0, 0, sizeof(double) * n * n);
    memset(d2, 0, sizeof(double) * n * n);
    for (i64 idet = 0; idet < ndet; ++idet) {
        u64 det = dets[idet];
        int no = occs_of(det, occs), nv = virs_of(det, n, virs);
        double val1 = coeffs[idet] * coeffs[idet];
        for (int i = 0; i < no; ++i) {
            i64 k = occs[i];
            d0[k * (n + 1)] += val1;
            for (int j = i + 1; j < no; ++j) {
                i64 l = occs[j];
                d2[n * k + l] += val1;
                d2[n * l + k] += val1;
            }
            for (int j = 0; j < nv; ++j) {
                i64 l = virs[j];
                u64 e = (det & ~(1ULL << k)) | (1ULL << l);
                i64 jdet = dset_find(&s, &e);
                if (jdet > idet) {
                    double val2 = coeffs[idet] * coeffs[jdet];
                    d0[n * k + l] += val2;
                    d0[n * l + k] += val2;
                }
            }
        }
    }
    dset_free(&s);
}

/* compute_rdms(FullCIWfn) (rdm.cpp:269-530): rdm1 (2,n,n) = aa, bb;
 * rdm2 (3,n,n,n,n) = aaaa, bbbb, abab with X[p,q,r,s] = <a+_p a+_q a_s a_r>. */
void oracle_rdms_fullci(i64 nbasis, i64 ndet, const u64 *dets, const double *coeffs, double *rdm1, double *rdm2) {
    dset s;
    dset_from(&s, 2, ndet, dets);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    double *aa = rdm1, *bb = rdm1 + n2, *aaaa = rdm2, *bbbb = rdm2 + n4, *abab = rdm2 + 2 * n4;
    memset(rdm1, 0, sizeof(double) * 2 * n2);
    memset(rdm2, 0, sizeof(double) * 3 * n4);
    i64 ou[64], od[64], vu[64], vd[64];
#define IX(p, q, r, s_) ((p) * n3 + (q) * n2 + (r) * n1 + (s_))
    for (i64 idet = 0; idet < ndet; ++idet) {
        u64 up = dets[2 * idet], dn = dets[2 * idet + 1];
        int nou = occs_of(up, ou), nod = occs_of(dn, od), nvu = virs_of(up, n1, vu), nvd = virs_of(dn, n1, vd);
        double val1 = coeffs[idet] * coeffs[idet], val2;
        for (int i = 0; i < nou; ++i) {
            i64 ii = ou[i];
            aa[(n1 + 1) * ii] += val1;
            for (int k = i + 1; k < nou; ++k) {
                i64 kk = ou[k];
                aaaa[IX(ii, kk, ii, kk)] += val1;
                aaaa[IX(ii, kk, kk, ii)] -= val1;
                aaaa[IX(kk, ii, ii, kk)] -= val1;
                aaaa[IX(kk, ii, kk, ii)] += val1;
            }
            for (int k = 0; k < nod; ++k) abab[IX(ii, od[k], ii, od[k])] += val1;
            for (int j = 0; j < nvu; ++j) {
                i64 jj = vu[j];
                u64 eu = (up & ~(1ULL << ii)) | (1ULL << jj);
                i64 sign_up = oracle_phase_single(up, ii, jj);
                u64 key[2] = {eu, dn};
                i64 jdet = dset_find(&s, key);
                if (jdet > idet) {
                    val2 = coeffs[idet] * coeffs[jdet] * sign_up;
                    aa[ii * n1 + jj] += val2;
                    aa[jj * n1 + ii] += val2;
                    for (int k = 0; k < nou; ++k) {
                        if (k == i) continue;
                        i64 kk = ou[k];
                        aaaa[IX(ii, kk, jj, kk)] += val2;
                        aaaa[IX(ii, kk, kk, jj)] -= val2;
                        aaaa[IX(kk, ii, kk, jj)] += val2;
                        aaaa[IX(kk, ii, jj, kk)] -= val2;
                        aaaa[IX(jj, kk, ii, kk)] += val2;
                        aaaa[IX(jj, kk, kk, ii)] -= val2;
                        aaaa[IX(kk, jj, ii, kk)] -= val2;
                        aaaa[IX(kk, jj, kk, ii)] += val2;
                    }
                    for (int k = 0; k < nod; ++k) {
                        i64 kk = od[k];
                        abab[IX(ii, kk, jj, kk)] += val2;
                        abab[IX(jj, kk, ii, kk)] += val2;
                    }
                }
                for (int k = 0; k < nod; ++k) {
                    i64 kk = od[k];
                    for (int l = 0; l < nvd; ++l) {
                        i64 ll = vd[l];
                        u64 k2[2] = {eu, (dn & ~(1ULL << kk)) | (1ULL << ll)};
                        jdet = dset_find(&s, k2);
                        if (jdet > idet) {
                            val2 = coeffs[idet] * coeffs[jdet] * sign_up * oracle_phase_single(dn, kk, ll);
                            abab[IX(ii, kk, jj, ll)] += val2;
                            abab[IX(jj, ll, ii, kk)] += val2;
                        }
                    }
                }
                for (int k = i + 1; k < nou; ++k) {
                    i64 kk = ou[k];
                    for (int l = j + 1; l < nvu; ++l) {
                        i64 ll = vu[l];
                        u64 k2[2] = {(eu & ~(1ULL << kk)) | (1ULL << ll), dn};
                        jdet = dset_find(&s, k2);
                        if (jdet > idet) {
                            val2 = coeffs[idet] * coeffs[jdet] * oracle_phase_double(up, ii, kk, jj, ll);
                            aaaa[IX(ii, kk, jj, ll)] += val2;
                            aaaa[IX(ii, kk, ll, jj)] -= val2;
                            aaaa[IX(kk, ii, jj, ll)] -= val2;
                            aaaa[IX(kk, ii, ll, jj)] += val2;
                            aaaa[IX(jj, ll, ii, kk)] += val2;
                            aaaa[IX(jj, ll, kk, ii)] -= val2;
                            aaaa[IX(ll, jj, ii, kk)] -= val2;
                            aaaa[IX(ll, jj, kk, ii)] += val2;
                        }
                    }
                }
            }
        }
        for (int i = 0; i < nod; ++i) {
            i64 ii = od[i];
            bb[(n1 + 1) * ii] += val1;
            for (int k = i + 1; k < nod; ++k) {
                i64 kk = od[k];
                bbbb[IX(ii, kk, ii, kk)] += val1;
                bbbb[IX(ii, kk, kk, ii)] -= val1;
                bbbb[IX(kk, ii, ii, kk)] -= val1;
                bbbb[IX(kk, ii, kk, ii)] += val1;
            }
            for (int j = 0; j < nvd; ++j) {
                i64 jj = vd[j];
                u64 ed = (dn & ~(1ULL << ii)) | (1ULL << jj);
                u64 key[2] = {up, ed};
                i64 jdet = dset_find(&s, key);
                if (jdet > idet) {
                    val2 = coeffs[idet] * coeffs[jdet] * oracle_phase_single(dn, ii, jj);
                    bb[ii * n1 + jj] += val2;
                    bb[jj * n1 + ii] += val2;
                    for (int k = 0; k < nou; ++k) {
                        i64 kk = ou[k];
                        abab[IX(kk, ii, kk, jj)] += val2;
                        abab[IX(kk, jj, kk, ii)] += val2;
                    }
                    for (int k = 0; k < nod; ++k) {
                        if (k == i) continue;
                        i64 kk = od[k];
                        bbbb[IX(ii, kk, jj, kk)] += val2;
                        bbbb[IX(ii, kk, kk, jj)] -= val2;
                        bbbb[IX(kk, ii, kk, jj)] += val2;
                        bbbb[IX(kk, ii, jj, kk)] -= val2;
                        bbbb[IX(jj, kk, ii, kk)] += val2;
                        bbbb[IX(jj, kk, kk, ii)] -= val2;
                        bbbb[IX(kk, jj, ii, kk)] -= val2;
                        bbbb[IX(kk, jj, kk, ii)] += val2;
                    }
                }
                for (int k = i + 1; k < nod; ++k) {
                    i64 kk = od[k];
                    for (int l = j + 1; l < nvd; ++l) {
                        i64 ll = vd[l];
                        u64 k2[2] = {up, (ed & ~(1ULL << kk)) | (1ULL << ll)};
                        jdet = dset_find(&s, k2);
                        if (jdet > idet) {
                            val2 = coeffs[idet] * coeffs[jdet] * oracle_phase_double(dn, ii, kk, jj, ll);
                            bbbb[IX(ii, kk, jj, ll)] += val2;
                            bbbb[IX(ii, kk, ll, jj)] -= val2;
                            bbbb[IX(kk, ii, jj, ll)] -= val2;
                            bbbb[IX(kk, ii, ll, jj)] += val2;
                            bbbb[IX(jj, ll, ii, kk)] += val2;
                            bbbb[IX(ll, jj, ii, kk)] -= val2;
                            bbbb[IX(jj, ll, kk, ii)] -= val2;
                            bbbb[IX(ll, jj, kk, ii)] += val2;
                        }
                    }
                }
            }
        }
    }
    dset_free(&s);
}

/* compute_rdms(GenCIWfn) by correct second-quantized algebra (the reference
 * routine at rdm.cpp:532-632 has index typos and writes out of bounds).
 * rdm1[p,q] = <a+_p a_q>, rdm2[p,q,r,s] = <a+_p a+_q a_s a_r>. Every ordered
 * ket J and every excitation J -> I with I in the wave function contributes
 * c_I c_J <I|op|J>, so both triangles are produced directly. */
void oracle_rdms_genci(i64 nbasis, i64 ndet, const u64 *dets, const double *coeffs, double *rdm1, double *rdm2) {
    dset s;
    dset_from(&s, 1, ndet, dets);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    memset(rdm1, 0, sizeof(double) * n2);
    memset(rdm2, 0, sizeof(double) * n4);
    i64 occs[64], virs[64];
    for (i64 jd = 0; jd < ndet; ++jd) {
        u64 det = dets[jd];
        int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
        double cj = coeffs[jd], c2 = cj * cj;
        for (int i = 0; i < no; ++i) {
            i64 p = occs[i];
            rdm1[p * n1 + p] += c2;
            for (int k = 0; k < no; ++k) {
                if (k == i) continue;
                i64 q = occs[k];
                rdm2[IX(p, q, p, q)] += c2;
                rdm2[IX(p, q, q, p)] -= c2;
            }
        }
        for (int i = 0; i < no; ++i) {
            i64 ii = occs[i];
            for (int j = 0; j < nv; ++j) {
                i64 aa = virs[j];
                u64 e = (det & ~(1ULL << ii)) | (1ULL << aa);
                i64 id = dset_find(&s, &e);
                if (id >= 0) {
                    /* <I| a+_a a_i |J> */
                    double v = coeffs[id] * cj * oracle_phase_single(det, ii, aa);
                    rdm1[aa * n1 + ii] += v;
                    for (int k = 0; k < no; ++k) {
                        if (k == i) continue;
                        i64 kk = occs[k];
                        rdm2[IX(aa, kk, ii, kk)] += v;
                        rdm2[IX(kk, aa, kk, ii)] += v;
                        rdm2[IX(aa, kk, kk, ii)] -= v;
                        rdm2[IX(kk, aa, ii, kk)] -= v;
                    }
                }
                for (int k = i + 1; k < no; ++k) {
                    i64 kk = occs[k];
                    for (int l = j + 1; l < nv; ++l) {
                        i64 ll = virs[l];
                        u64 e2 = (e & ~(1ULL << kk)) | (1ULL << ll);
                        id = dset_find(&s, &e2);
                        if (id >= 0) {
                            /* <I| a+_a a+_l a_k a_i |J> = phase_double(i, k, a, l) */
                            double v = coeffs[id] * cj * oracle_phase_double(det, ii, kk, aa, ll);
                            rdm2[IX(aa, ll, ii, kk)] += v;
                            rdm2[IX(aa, ll, kk, ii)] -= v;
                            rdm2[IX(ll, aa, ii, kk)] -= v;
                            rdm2[IX(ll, aa, kk, ii)] += v;
                        }
                    }
                }
            }
        }
    }
#undef IX
    dset_free(&s);
}

/* Diagonal of the CI matrix for every determinant (the diagonal element of
 * add_row, sparseop.cpp:283-292/367-372 for FullCI, :228-236 DOCI, :443-451 GenCI). */
void oracle_diagonal(int kind, i64 nbasis, i64 nup, i64 ndn, i64 ndet, const u64 *dets, const double *one_mo,
                     const double *two_mo, double *diag) {
    for (i64 i = 0; i < ndet; ++i) {
        /* A one-row operator restricted to the diagonal: build with a one-det set. */
        int nkey = (kind == KIND_FULLCI) ? 2 : 1;
        oracle_csr *c = oracle_csr_build(kind, nbasis, nup, ndn, 1, dets + i * nkey, one_mo, two_mo, 0, 1, 1, 1);
        diag[i] = c->data[c->nnz - 1];
        oracle_csr_free(c);
    }
}

/* ------------------------------------------------------------------------ */
/* Threaded symmetric product for config-sized golden vectors               */
/* ------------------------------------------------------------------------ */

/* perform_op_symm (sparseop.cpp:148-159) over a lower+diagonal CSR held as
 * int64 row pointers, int32 columns and float64 values (nrow rows, row i global
 * row i): y = (L + L^T - D) x. The rows are split into nthread ranges of equal
 * nonzero count; each range accumulates into a private y (its own rows and the
 * transposed half), and the private vectors are summed in range order, so the
 * result does not depend on thread timing. ecore is not added. */
#include <pthread.h>

typedef struct {
    i64 r0, r1, n;
    const i64 *indptr;
    const int32_t *indices;
    const double *data, *x;
    double *y;
} symv_job;

static void *symv_worker(void *arg) {
    symv_job *J = (symv_job *)arg;
    for (i64 i = J->r0; i < J->r1; ++i) {
        double acc = 0.0, xi = J->x[i];
        for (i64 p = J->indptr[i]; p < J->indptr[i + 1]; ++p) {
            i64 j = J->indices[p];
            acc += J->data[p] * J->x[j];
            if (j != i) J->y[j] += J->data[p] * xi;
        }
        J->y[i] += acc;
    }
    return NULL;
}

int oracle_symv_lower_mt(i64 nrow, const i64 *indptr, const int32_t *indices, const double *data, const double *x,
                         double *y, int nthread) {
    if (nthread < 1) nthread = 1;
    if (nthread > 64) nthread = 64;
    symv_job jobs[64];
    pthread_t th[64];
    i64 nnz = indptr[nrow], r = 0;
    for (int t = 0; t < nthread; ++t) {
        i64 target = nnz / nthread * (t + 1), r1 = r;
        if (t == nthread - 1) r1 = nrow;
        else while (r1 < nrow && indptr[r1] < target) ++r1;
        jobs[t].r0 = r;
        jobs[t].r1 = r1;
        jobs[t].n = nrow;
        jobs[t].indptr = indptr;
        jobs[t].indices = indices;
        jobs[t].data = data;
        jobs[t].x = x;
        jobs[t].y = (double *)calloc((size_t)nrow, sizeof(double));
        if (!jobs[t].y) return -1;
        r = r1;
    }
    for (int t = 0; t < nthread; ++t) pthread_create(&th[t], NULL, symv_worker, &jobs[t]);
    for (int t = 0; t < nthread; ++t) pthread_join(th[t], NULL);
    for (i64 i = 0; i < nrow; ++i) {
        double s = 0.0;
        for (int t = 0; t < nthread; ++t) s += jobs[t].y[i];
        y[i] = s;
    }
    for (int t = 0; t < nthread; ++t) free(jobs[t].y);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Epstein-Nesbet PT2 (enpt2.cpp:76-235 FullCI, :257-328 GenCI, :342-370)    */
/* ------------------------------------------------------------------------ */

/* Diagonal element of one determinant (compute_enpt2_thread_gather,
 * enpt2.cpp:37-70 FullCI, :237-255 GenCI): one_mo[j,j], same-spin pairs
 * <jl|jl> - <jl|lj>, opposite-spin pairs <jl|jl>, in the reference's order. */
static double pt2_diag(int kind, const hamv *H, const u64 *k) {
    i64 n1 = H->n1, n2 = n1 * n1, n3 = n2 * n1;
    i64 ou[64], od[64];
    int nou = occs_of(k[0], ou), nod = (kind == KIND_FULLCI) ? occs_of(k[1], od) : 0;
    double diag = 0.0;
    for (int i = 0; i < nou; ++i) {
        i64 j = ou[i];
        diag += H->one_mo[(n1 + 1) * j];
        for (int q = i + 1; q < nou; ++q) {
            i64 l = ou[q];
            diag += G(j, l, j, l) - G(j, l, l, j);
        }
        for (int q = 0; q < nod; ++q) {
            i64 l = od[q];
            diag += G(j, l, j, l);
        }
    }
    for (int i = 0; i < nod; ++i) {
        i64 j = od[i];
        diag += H->one_mo[(n1 + 1) * j];
        for (int q = i + 1; q < nod; ++q) {
            i64 l = od[q];
            diag += G(j, l, j, l) - G(j, l, l, j);
        }
    }
    return diag;
}

typedef struct {
    dset t;        /* external determinants in first-occurrence order */
    double *num;   /* sum_i H_ai c_i */
    double *diag;  /* H_aa */
    i64 cap;
} pt2_terms;

static void pt2_add(pt2_terms *P, int kind, const hamv *H, const dset *s, const u64 *k, double v) {
    if (dset_find(s, k) != -1) return; /* index_det_from_rank(rank) == -1 only */
    i64 a = dset_find(&P->t, k);
    if (a < 0) {
        a = dset_add(&P->t, k);
        if (a >= P->cap) {
            P->cap = 2 * (a + 1);
            P->num = (double *)realloc(P->num, sizeof(double) * P->cap);
            P->diag = (double *)realloc(P->diag, sizeof(double) * P->cap);
        }
        P->num[a] = 0.0;
        P->diag[a] = pt2_diag(kind, H, k);
    }
    P->num[a] += v;
}

/* compute_enpt2 (enpt2.cpp:342-370) without the final "energy +": returns
 * sum_a (sum_i H_ai c_i)^2 / (e - H_aa) over the determinants a outside the set
 * reached from some i by an excitation with |H_ai| > eps/|c_i| (e = energy -
 * ecore). FullCI (kind 1) and GenCI (kind 2); a DOCI set enters as the FullCI set
 * of its (d, d) determinants (enpt2.cpp:382-386). The contributions of one
 * external determinant are summed in serial order (source ascending, reference
 * loop order); the reference sums per thread and then merges. */
double oracle_enpt2(int kind, i64 nbasis, i64 nup, i64 ndn, i64 ndet, const u64 *dets, const double *one_mo,
                    const double *two_mo, const double *coeffs, double e, double eps, i64 *nterms) {
    int nkey = (kind == KIND_FULLCI) ? 2 : 1;
    dset s;
    dset_from(&s, nkey, ndet, dets);
    hamv H;
    ham_init(&H, kind, nbasis, nup, ndn, one_mo, two_mo);
    const hamv *HH = &H;
    pt2_terms P;
    dset_init(&P.t, nkey, 1024);
    P.cap = 1024;
    P.num = (double *)malloc(sizeof(double) * P.cap);
    P.diag = (double *)malloc(sizeof(double) * P.cap);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1;
    i64 occs[64], virs[64], ou[64], od[64], vu[64], vd[64];
#define GG(p, q, r, s_) (HH->two_mo[(p) * n3 + (q) * n2 + (r) * n1 + (s_)])
    for (i64 idet = 0; idet < ndet; ++idet) {
        const double c = coeffs[idet], eps_i = eps / fabs(c);
        if (kind == KIND_GENCI) {
            u64 det = dets[idet];
            int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
            for (int i = 0; i < no; ++i) {
                i64 ii = occs[i];
                for (int j = 0; j < nv; ++j) {
                    i64 jj = virs[j];
                    u64 ex = (det & ~(1ULL << ii)) | (1ULL << jj);
                    double val = one_mo[n1 * ii + jj];
                    for (int k = 0; k < no; ++k) val += GG(ii, occs[k], jj, occs[k]) - GG(ii, occs[k], occs[k], jj);
                    if (fabs(val) > eps_i) pt2_add(&P, kind, HH, &s, &ex, val * oracle_phase_single(det, ii, jj) * c);
                    for (int k = i + 1; k < no; ++k)
                        for (int l = j + 1; l < nv; ++l) {
                            i64 kk = occs[k], ll = virs[l];
                            val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                            if (fabs(val) > eps_i) {
                                u64 e2 = (ex & ~(1ULL << kk)) | (1ULL << ll);
                                pt2_add(&P, kind, HH, &s, &e2, val * oracle_phase_double(det, ii, kk, jj, ll) * c);
                            }
                        }
                }
            }
            continue;
        }
        u64 up = dets[2 * idet], dn = dets[2 * idet + 1];
        int nou = occs_of(up, ou), nod = occs_of(dn, od), nvu = virs_of(up, n1, vu), nvd = virs_of(dn, n1, vd);
        for (int i = 0; i < nou; ++i) {
            i64 ii = ou[i];
            for (int j = 0; j < nvu; ++j) {
                i64 jj = vu[j];
                u64 eu = (up & ~(1ULL << ii)) | (1ULL << jj);
                i64 sign_up = oracle_phase_single(up, ii, jj);
                double val = one_mo[n1 * ii + jj];
                for (int k = 0; k < nou; ++k) val += GG(ii, ou[k], jj, ou[k]) - GG(ii, ou[k], ou[k], jj);
                for (int k = 0; k < nod; ++k) val += GG(ii, od[k], jj, od[k]);
                if (fabs(val) > eps_i) {
                    u64 key[2] = {eu, dn};
                    pt2_add(&P, kind, HH, &s, key, val * sign_up * c);
                }
                for (int k = 0; k < nod; ++k)
                    for (int l = 0; l < nvd; ++l) {
                        i64 kk = od[k], ll = vd[l];
                        val = GG(ii, kk, jj, ll);
                        if (fabs(val) > eps_i) {
                            u64 key[2] = {eu, (dn & ~(1ULL << kk)) | (1ULL << ll)};
                            pt2_add(&P, kind, HH, &s, key, val * (sign_up * oracle_phase_single(dn, kk, ll)) * c);
                        }
                    }
                for (int k = i + 1; k < nou; ++k)
                    for (int l = j + 1; l < nvu; ++l) {
                        i64 kk = ou[k], ll = vu[l];
                        val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                        if (fabs(val) > eps_i) {
                            u64 key[2] = {(eu & ~(1ULL << kk)) | (1ULL << ll), dn};
                            pt2_add(&P, kind, HH, &s, key, val * oracle_phase_double(up, ii, kk, jj, ll) * c);
                        }
                    }
            }
        }
        for (int i = 0; i < nod; ++i) {
            i64 ii = od[i];
            for (int j = 0; j < nvd; ++j) {
                i64 jj = vd[j];
                u64 ed = (dn & ~(1ULL << ii)) | (1ULL << jj);
                double val = one_mo[n1 * ii + jj];
                for (int k = 0; k < nou; ++k) val += GG(ii, ou[k], jj, ou[k]);
                for (int k = 0; k < nod; ++k) val += GG(ii, od[k], jj, od[k]) - GG(ii, od[k], od[k], jj);
                if (fabs(val) > eps_i) {
                    u64 key[2] = {up, ed};
                    pt2_add(&P, kind, HH, &s, key, val * oracle_phase_single(dn, ii, jj) * c);
                }
                for (int k = i + 1; k < nod; ++k)
                    for (int l = j + 1; l < nvd; ++l) {
                        i64 kk = od[k], ll = vd[l];
                        val = GG(ii, kk, jj, ll) - GG(ii, kk, ll, jj);
                        if (fabs(val) > eps_i) {
                            u64 key[2] = {up, (ed & ~(1ULL << kk)) | (1ULL << ll)};
                            pt2_add(&P, kind, HH, &s, key, val * oracle_phase_double(dn, ii, kk, jj, ll) * c);
                        }
                    }
            }
        }
    }
#undef GG
    double corr = 0.0;
    for (i64 a = 0; a < P.t.n; ++a) corr += P.num[a] * P.num[a] / (e - P.diag[a]);
    if (nterms) *nterms = P.t.n;
    free(P.num);
    free(P.diag);
    dset_free(&P.t);
    ham_free(&H);
    dset_free(&s);
    return corr;
}

/* ------------------------------------------------------------------------ */
/* Transition RDMs (rdm.cpp:634-899; GenCI by correct math)                  */
/* ------------------------------------------------------------------------ */

/* compute_transition_rdms(DOCIWfn) (rdm.cpp:634-672): for every det I of set 1
 * and its pair hops J (k -> l) present in set 2, d0[k,l] += c1_I c2_J; the
 * diagonal I itself (if in set 2) adds c1_I c2_I' to d0[k,k] and d2[k,l], d2[l,k]
 * over its occupied pairs. */
void oracle_trdms_doci(i64 nbasis, i64 ndet1, const u64 *dets1, i64 ndet2, const u64 *dets2, const double *c1,
                       const double *c2, double *d0, double *d2) {
    dset s;
    dset_from(&s, 1, ndet2, dets2);
    i64 n = nbasis, occs[64], virs[64];
    memset(d0, 0, sizeof(double) * n * n);
    memset(d2, 0, sizeof(double) * n * n);
    for (i64 idet = 0; idet < ndet1; ++idet) {
        u64 det = dets1[idet];
        int no = occs_of(det, occs), nv = virs_of(det, n, virs);
        i64 jd = dset_find(&s, &det);
        double val1 = (jd != -1) ? c1[idet] * c2[jd] : 0.0;
        for (int i = 0; i < no; ++i) {
            i64 k = occs[i];
            d0[k * (n + 1)] += val1;
            for (int j = i + 1; j < no; ++j) {
                i64 l = occs[j];
                d2[n * k + l] += val1;
                d2[n * l + k] += val1;
            }
            for (int j = 0; j < nv; ++j) {
                i64 l = virs[j];
                u64 e = (det & ~(1ULL << k)) | (1ULL << l);
                jd = dset_find(&s, &e);
                if (jd != -1) d0[n * k + l] += c1[idet] * c2[jd];
            }
        }
    }
    dset_free(&s);
}

/* compute_transition_rdms(FullCIWfn) (rdm.cpp:674-899): rdm1 = (aa, bb), rdm2 =
 * (aaaa, bbbb, abab); element [p,q] = <1| a+_p a_q |2>, [p,q,r,s] =
 * <1| a+_p a+_q a_s a_r |2>, one term per (I in set 1, J in set 2) with the
 * reference's index images. */
void oracle_trdms_fullci(i64 nbasis, i64 ndet1, const u64 *dets1, i64 ndet2, const u64 *dets2, const double *c1,
                         const double *c2, double *rdm1, double *rdm2) {
    dset s;
    dset_from(&s, 2, ndet2, dets2);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    double *aa = rdm1, *bb = rdm1 + n2, *aaaa = rdm2, *bbbb = rdm2 + n4, *abab = rdm2 + 2 * n4;
    memset(rdm1, 0, sizeof(double) * 2 * n2);
    memset(rdm2, 0, sizeof(double) * 3 * n4);
    i64 ou[64], od[64], vu[64], vd[64];
#define A4(X, p, q, r, t) X[(p) * n3 + (q) * n2 + (r) * n1 + (t)]
    for (i64 idet = 0; idet < ndet1; ++idet) {
        u64 up = dets1[2 * idet], dn = dets1[2 * idet + 1];
        int nou = occs_of(up, ou), nod = occs_of(dn, od), nvu = virs_of(up, n1, vu), nvd = virs_of(dn, n1, vd);
        u64 key[2] = {up, dn};
        i64 jd = dset_find(&s, key);
        const double ci = c1[idet];
        double val1 = (jd != -1) ? ci * c2[jd] : 0.0;
        for (int i = 0; i < nou; ++i) {
            i64 ii = ou[i];
            aa[(n1 + 1) * ii] += val1;
            for (int k = i + 1; k < nou; ++k) {
                i64 kk = ou[k];
                A4(aaaa, ii, kk, ii, kk) += val1;
                A4(aaaa, ii, kk, kk, ii) -= val1;
                A4(aaaa, kk, ii, ii, kk) -= val1;
                A4(aaaa, kk, ii, kk, ii) += val1;
            }
            for (int k = 0; k < nod; ++k) A4(abab, ii, od[k], ii, od[k]) += val1;
            for (int j = 0; j < nvu; ++j) {
                i64 jj = vu[j];
                u64 eu = (up & ~(1ULL << ii)) | (1ULL << jj);
                i64 sign_up = oracle_phase_single(up, ii, jj);
                u64 k1[2] = {eu, dn};
                jd = dset_find(&s, k1);
                if (jd != -1) {
                    double v = ci * c2[jd] * sign_up;
                    aa[ii * n1 + jj] += v;
                    for (int k = 0; k < nou; ++k) {
                        if (k == i) continue;
                        i64 kk = ou[k];
                        A4(aaaa, ii, kk, jj, kk) += v;
                        A4(aaaa, ii, kk, kk, jj) -= v;
                        A4(aaaa, kk, ii, kk, jj) += v;
                        A4(aaaa, kk, ii, jj, kk) -= v;
                    }
                    for (int k = 0; k < nod; ++k) A4(abab, ii, od[k], jj, od[k]) += v;
                }
                for (int k = 0; k < nod; ++k)
                    for (int l = 0; l < nvd; ++l) {
                        i64 kk = od[k], ll = vd[l];
                        u64 k2[2] = {eu, (dn & ~(1ULL << kk)) | (1ULL << ll)};
                        jd = dset_find(&s, k2);
                        if (jd != -1) A4(abab, ii, kk, jj, ll) += ci * c2[jd] * sign_up * oracle_phase_single(dn, kk, ll);
                    }
                for (int k = i + 1; k < nou; ++k)
                    for (int l = j + 1; l < nvu; ++l) {
                        i64 kk = ou[k], ll = vu[l];
                        u64 k2[2] = {(eu & ~(1ULL << kk)) | (1ULL << ll), dn};
                        jd = dset_find(&s, k2);
                        if (jd != -1) {
                            double v = ci * c2[jd] * oracle_phase_double(up, ii, kk, jj, ll);
                            A4(aaaa, ii, kk, jj, ll) += v;
                            A4(aaaa, ii, kk, ll, jj) -= v;
                            A4(aaaa, kk, ii, jj, ll) -= v;
                            A4(aaaa, kk, ii, ll, jj) += v;
                        }
                    }
            }
        }
        for (int i = 0; i < nod; ++i) {
            i64 ii = od[i];
            bb[(n1 + 1) * ii] += val1;
            for (int k = i + 1; k < nod; ++k) {
                i64 kk = od[k];
                A4(bbbb, ii, kk, ii, kk) += val1;
                A4(bbbb, ii, kk, kk, ii) -= val1;
                A4(bbbb, kk, ii, ii, kk) -= val1;
                A4(bbbb, kk, ii, kk, ii) += val1;
            }
            for (int j = 0; j < nvd; ++j) {
                i64 jj = vd[j];
                u64 ed = (dn & ~(1ULL << ii)) | (1ULL << jj);
                u64 k1[2] = {up, ed};
                jd = dset_find(&s, k1);
                if (jd != -1) {
                    double v = ci * c2[jd] * oracle_phase_single(dn, ii, jj);
                    bb[ii * n1 + jj] += v;
                    for (int k = 0; k < nou; ++k) A4(abab, ou[k], ii, ou[k], jj) += v;
                    for (int k = 0; k < nod; ++k) {
                        if (k == i) continue;
                        i64 kk = od[k];
                        A4(bbbb, ii, kk, jj, kk) += v;
                        A4(bbbb, ii, kk, kk, jj) -= v;
                        A4(bbbb, kk, ii, kk, jj) += v;
                        A4(bbbb, kk, ii, jj, kk) -= v;
                    }
                }
                for (int k = i + 1; k < nod; ++k)
                    for (int l = j + 1; l < nvd; ++l) {
                        i64 kk = od[k], ll = vd[l];
                        u64 k2[2] = {up, (ed & ~(1ULL << kk)) | (1ULL << ll)};
                        jd = dset_find(&s, k2);
                        if (jd != -1) {
                            double v = ci * c2[jd] * oracle_phase_double(dn, ii, kk, jj, ll);
                            A4(bbbb, ii, kk, jj, ll) += v;
                            A4(bbbb, ii, kk, ll, jj) -= v;
                            A4(bbbb, kk, ii, jj, ll) -= v;
                            A4(bbbb, kk, ii, ll, jj) += v;
                        }
                    }
            }
        }
    }
#undef A4
    dset_free(&s);
}

/* Transition RDMs of two GenCI sets by the operator algebra (the reference
 * routine, rdm.cpp:901-1009, reads coeffs2 at the wrong index and writes past
 * its arrays): rdm1[p,q] = <1| a+_p a_q |2>, rdm2[p,q,r,s] = <1| a+_p a+_q a_s a_r |2>. */
void oracle_trdms_genci(i64 nbasis, i64 ndet1, const u64 *dets1, i64 ndet2, const u64 *dets2, const double *c1,
                        const double *c2, double *rdm1, double *rdm2) {
    dset s;
    dset_from(&s, 1, ndet2, dets2);
    i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1, n4 = n2 * n2;
    memset(rdm1, 0, sizeof(double) * n2);
    memset(rdm2, 0, sizeof(double) * n4);
    i64 occs[64], virs[64];
#define A4(p, q, r, t) rdm2[(p) * n3 + (q) * n2 + (r) * n1 + (t)]
    for (i64 idet = 0; idet < ndet1; ++idet) {
        u64 det = dets1[idet];
        int no = occs_of(det, occs), nv = virs_of(det, n1, virs);
        i64 jd = dset_find(&s, &det);
        const double ci = c1[idet];
        double val1 = (jd != -1) ? ci * c2[jd] : 0.0;
        for (int i = 0; i < no; ++i) {
            i64 ii = occs[i];
            rdm1[(n1 + 1) * ii] += val1;
            for (int k = i + 1; k < no; ++k) {
                i64 kk = occs[k];
                A4(ii, kk, ii, kk) += val1;
                A4(ii, kk, kk, ii) -= val1;
                A4(kk, ii, ii, kk) -= val1;
                A4(kk, ii, kk, ii) += val1;
            }
            for (int j = 0; j < nv; ++j) {
                i64 jj = virs[j];
                u64 e = (det & ~(1ULL << ii)) | (1ULL << jj);
                jd = dset_find(&s, &e);
                if (jd != -1) {
                    double v = ci * c2[jd] * oracle_phase_single(det, ii, jj);
                    rdm1[ii * n1 + jj] += v;
                    for (int k = 0; k < no; ++k) {
                        if (k == i) continue;
                        i64 kk = occs[k];
                        A4(ii, kk, jj, kk) += v;
                        A4(ii, kk, kk, jj) -= v;
                        A4(kk, ii, kk, jj) += v;
                        A4(kk, ii, jj, kk) -= v;
                    }
                }
                for (int k = i + 1; k < no; ++k)
                    for (int l = j + 1; l < nv; ++l) {
                        i64 kk = occs[k], ll = virs[l];
                        u64 e2 = (e & ~(1ULL << kk)) | (1ULL << ll);
                        jd = dset_find(&s, &e2);
                        if (jd != -1) {
                            double v = ci * c2[jd] * oracle_phase_double(det, ii, kk, jj, ll);
                            A4(ii, kk, jj, ll) += v;
                            A4(ii, kk, ll, jj) -= v;
                            A4(kk, ii, jj, ll) -= v;
                            A4(kk, ii, ll, jj) += v;
                        }
                    }
            }
        }
    }
#undef A4
    dset_free(&s);
}
