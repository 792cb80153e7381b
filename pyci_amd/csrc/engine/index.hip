// Determinant index on the GPU, the colex fill of the full space, and the
// device-wide exclusive scan used by the CSR builder.
//
// Reference counterparts: Wfn::dict (SpookyHash128 + phmap, onespinwfn.cpp:123-143,
// twospinwfn.cpp:129-147) and add_all_dets (onespinwfn.cpp:173-217,
// twospinwfn.cpp:181-245). Here the index of a set that equals the full space in
// colex order is the perfect rank r_up * C(nb, n_dn) + r_dn (no memory probe);
// any other set gets an open-addressing table keyed by the determinant words.
#include "internal.hpp"

namespace pe {

i64 binom_host(i64 n, i64 k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 b = 1;
    for (i64 d = 1; d <= k; ++d) b = b * (unsigned __int128)(n - k + d) / (unsigned __int128)d;
    return (i64)b;
}

WfnView view_of(const pyci_wfn *w) {
    WfnView v;
    v.kind = w->kind;
    v.nkey = w->nkey;
    v.nb = (int)w->nbasis;
    v.nup = (int)w->nup;
    v.ndn = (int)w->ndn;
    v.ndet = w->ndet;
    v.dets = w->dets.p;
    v.colex = w->colex ? 1 : 0;
    v.nsd = w->nsd;
    v.binom = w->ctx->binom.p;
    v.hkeys = w->hkeys.p;
    v.hvals = w->hvals.p;
    v.htags = w->htags.p;
    v.hfilt = w->hfilt.p;
    v.fmask = w->fbits ? (u64)(w->fbits - 1) : 0;
    v.hmask = w->hcap ? (u64)(w->hcap - 1) : 0;
    return v;
}

HamView view_of(const pyci_ham *h) {
    HamView v;
    v.n = (int)h->nbasis;
    v.one = h->one_mo.p;
    v.two = h->two_mo.p;
    return v;
}

// ---------------------------------------------------------------------------
// Colex check: flags[0] stays 0 iff det i has colex (perfect) rank i for all i.
// ---------------------------------------------------------------------------
__global__ void k_colex_check(WfnView w, int *bad) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.ndet) return;
    u64 m = full_mask(w.nb);
    bool ok;
    if (w.nkey == 2) {
        u64 a = w.dets[2 * i], b = w.dets[2 * i + 1];
        ok = (a & ~m) == 0 && (b & ~m) == 0 && __popcll(a) == w.nup && __popcll(b) == w.ndn &&
             colex_rank(a, w.binom) * w.nsd + colex_rank(b, w.binom) == i;
    } else {
        u64 a = w.dets[i];
        ok = (a & ~m) == 0 && __popcll(a) == w.nup && colex_rank(a, w.binom) == i;
    }
    if (!ok) atomicOr(bad, 1);
}

__global__ void k_hash_clear(i64 *vals, u64 *tags, i64 n) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vals[i] = -1;
    if (i < n / 8) tags[i] = 0;
}

// Inserts dets[first..ndet) with value = position. Dets are distinct, so a slot
// is claimed by CAS on its value and the key is written by the claimer; no
// lookups run concurrently with this kernel.
__global__ void k_hash_insert(const u64 *dets, int nkey, i64 first, i64 ndet, u64 *keys, i64 *vals, u64 *tags,
                              u64 mask, unsigned *filt, u64 fmask) {
    i64 i = first + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndet) return;
    u64 a = dets[nkey * i], b = (nkey == 2) ? dets[nkey * i + 1] : 0ULL;
    const u64 hk = hash_key(a, b);
    if (filt) {
        const u64 f = (hk >> 24) & fmask;
        atomicOr(&filt[f >> 5], 1u << (f & 31));
    }
    u64 h = hk & mask;
    for (;;) {
        unsigned long long old = atomicCAS((unsigned long long *)&vals[h], (unsigned long long)(-1LL),
                                           (unsigned long long)i);
        if (old == (unsigned long long)(-1LL)) {
            keys[2 * h] = a;
            keys[2 * h + 1] = b;
            reinterpret_cast<unsigned char *>(tags)[h] = (unsigned char)((hk >> 57) | 0x80ULL);
            return;
        }
        h = (h + 1) & mask;
    }
}

__global__ void k_lookup(WfnView w, const u64 *q, i64 n, i64 *out) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 a = q[w.nkey * i], b = (w.nkey == 2) ? q[w.nkey * i + 1] : 0ULL;
    u64 m = full_mask(w.nb);
    if ((a & ~m) || (b & ~m) || __popcll(a) != w.nup || (w.nkey == 2 && __popcll(b) != w.ndn)) {
        out[i] = -1;
        return;
    }
    i64 r = wfn_lookup(w, a, b);
    out[i] = (r >= 0 && r < w.ndet) ? r : -1;
}

void wfn_build_index(pyci_wfn *w) {
    pyci_ctx *ctx = w->ctx;
    hipStream_t s = ctx->stream;
    // 1) is it the full space in colex order?
    i64 total = binom_host(w->nbasis, w->nup) * (w->nkey == 2 ? binom_host(w->nbasis, w->ndn) : 1);
    w->colex = false;
    if (w->ndet == total && w->ndet > 0) {
        DevBuf<int> bad(1);
        PE_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        w->colex = true;  // view uses colex ranks during the check
        WfnView v = view_of(w);
        k_colex_check<<<grid_for(w->ndet, 256), 256, 0, s>>>(v, bad.p);
        PE_LAUNCH_CHECK();
        int hb = 0;
        PE_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PE_HIP(hipStreamSynchronize(s));
        w->colex = (hb == 0);
    }
    if (w->colex) {
        w->hkeys.release();
        w->hvals.release();
        w->htags.release();
        w->hfilt.release();
        w->hcap = 0;
        w->fbits = 0;
        return;
    }
    // 2) hash table at load factor <= 1/2
    i64 cap = 1024;
    while (cap < 2 * w->ndet + 2) cap <<= 1;
    w->hcap = cap;
    w->hkeys.alloc((size_t)(2 * cap));
    w->hvals.alloc((size_t)cap);
    w->htags.alloc((size_t)(cap / 8));
    k_hash_clear<<<grid_for(cap, 256), 256, 0, s>>>(w->hvals.p, w->htags.p, cap);
    PE_LAUNCH_CHECK();
    // the presence filter once the tag array outgrows an XCD's 4 MB L2: 2 bits per slot,
    // at least 4 per determinant (the table's load is at most 1/2), 1/8 the tags' bytes
    w->hfilt.release();
    w->fbits = 0;
    if (cap >= ((i64)1 << 22)) {
        w->fbits = 2 * cap;
        w->hfilt.alloc((size_t)(w->fbits / 32));
        PE_HIP(hipMemsetAsync(w->hfilt.p, 0, sizeof(unsigned) * (size_t)(w->fbits / 32), s));
    }
    if (w->ndet > 0) {
        k_hash_insert<<<grid_for(w->ndet, 256), 256, 0, s>>>(w->dets.p, w->nkey, 0, w->ndet, w->hkeys.p, w->hvals.p,
                                                              w->htags.p, (u64)(cap - 1), w->hfilt.p,
                                                              w->fbits ? (u64)(w->fbits - 1) : 0);
        PE_LAUNCH_CHECK();
    }
}

// Appends to the hash index (or rebuilds when the table must grow).
void wfn_index_append(pyci_wfn *w, i64 first) {
    if (w->colex || w->hcap == 0 || 2 * w->ndet + 2 > w->hcap) {
        w->colex = false;
        wfn_build_index(w);
        return;
    }
    hipStream_t s = w->ctx->stream;
    i64 n = w->ndet - first;
    if (n > 0) {
        k_hash_insert<<<grid_for(n, 256), 256, 0, s>>>(w->dets.p, w->nkey, first, w->ndet, w->hkeys.p, w->hvals.p,
                                                       w->htags.p, (u64)(w->hcap - 1), w->hfilt.p,
                                                       w->fbits ? (u64)(w->fbits - 1) : 0);
        PE_LAUNCH_CHECK();
    }
}

void wfn_lookup_batch(const pyci_wfn *w, const u64 *d_q, i64 n, i64 *d_out) {
    if (n <= 0) return;
    WfnView v = view_of(w);
    k_lookup<<<grid_for(n, 256), 256, 0, w->ctx->stream>>>(v, d_q, n, d_out);
    PE_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Colex fill of the full space (any nbasis; nword words per spin string).
// ---------------------------------------------------------------------------
__device__ u64 dev_binom(i64 n, i64 k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    u64 b = 1;
    for (i64 d = 1; d <= k; ++d) b = b * (u64)(n - k + d) / (u64)d;
    return b;
}

// strings[r * nword ...] = colex-unranked string r (common.cpp:141-153 semantics).
__global__ void k_unrank_strings(i64 nb, i64 nocc, i64 count, int nword, u64 *strings) {
    i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    u64 *out = strings + r * nword;
    for (int w = 0; w < nword; ++w) out[w] = 0;
    i64 rank = r, n = nb;
    for (i64 k = nocc; k >= 1; --k) {
        i64 p = n - 1;
        while ((i64)dev_binom(p, k) > rank) --p;
        out[p / 64] |= 1ULL << (p % 64);
        rank -= (i64)dev_binom(p, k);
        n = p;
    }
}

__global__ void k_fill_two_spin(const u64 *su, const u64 *sd, i64 nsu, i64 nsd, int nword, u64 *out) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsu * nsd) return;
    i64 iu = i / nsd, id = i % nsd;
    u64 *o = out + i * 2 * nword;
    for (int w = 0; w < nword; ++w) {
        o[w] = su[iu * nword + w];
        o[nword + w] = sd[id * nword + w];
    }
}

void fill_all_dets_device(pyci_ctx *ctx, int kind, i64 nb, i64 nup, i64 ndn, u64 *d_out, int nword) {
    hipStream_t s = ctx->stream;
    i64 nsu = binom_host(nb, nup);
    if (kind != PYCI_WFN_FULLCI) {
        if (nsu > 0) {
            k_unrank_strings<<<grid_for(nsu, 256), 256, 0, s>>>(nb, nup, nsu, nword, d_out);
            PE_LAUNCH_CHECK();
        }
        return;
    }
    i64 nsd = binom_host(nb, ndn);
    DevBuf<u64> su((size_t)(nsu * nword)), sd((size_t)(nsd * nword));
    k_unrank_strings<<<grid_for(nsu, 256), 256, 0, s>>>(nb, nup, nsu, nword, su.p);
    PE_LAUNCH_CHECK();
    k_unrank_strings<<<grid_for(nsd, 256), 256, 0, s>>>(nb, ndn, nsd, nword, sd.p);
    PE_LAUNCH_CHECK();
    if (nsu * nsd > 0) {
        k_fill_two_spin<<<grid_for(nsu * nsd, 256), 256, 0, s>>>(su.p, sd.p, nsu, nsd, nword, d_out);
        PE_LAUNCH_CHECK();
    }
    PE_HIP(hipStreamSynchronize(s));  // su/sd are freed on return
}

// ---------------------------------------------------------------------------
// Exclusive scan of int64 (block-local scan + host scan of block totals).
// ---------------------------------------------------------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums(const i64 *d, i64 n, i64 *sums) {
    __shared__ i64 red[SCAN_THREADS];
    i64 base = (i64)blockIdx.x * SCAN_TILE;
    i64 acc = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        i64 i = base + (i64)t * SCAN_THREADS + threadIdx.x;
        if (i < n) acc += d[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = SCAN_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_scan(i64 *d, i64 n, const i64 *offsets) {
    __shared__ i64 part[SCAN_THREADS];
    i64 base = (i64)blockIdx.x * SCAN_TILE + (i64)threadIdx.x * SCAN_ITEMS;
    i64 v[SCAN_ITEMS];
    i64 acc = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        i64 i = base + t;
        v[t] = (i < n) ? d[i] : 0;
        acc += v[t];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    // Hillis-Steele inclusive scan over thread totals
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        i64 x = ((int)threadIdx.x >= o) ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    i64 run = offsets[blockIdx.x] + part[threadIdx.x] - acc;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        i64 i = base + t;
        if (i < n) d[i] = run;
        run += v[t];
    }
}

i64 exclusive_scan(pyci_ctx *ctx, i64 *d_data, i64 n) {
    if (n <= 0) return 0;
    hipStream_t s = ctx->stream;
    i64 ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf<i64> sums((size_t)ntiles);
    k_tile_sums<<<(unsigned)ntiles, SCAN_THREADS, 0, s>>>(d_data, n, sums.p);
    PE_LAUNCH_CHECK();
    std::vector<i64> h((size_t)ntiles);
    PE_HIP(hipMemcpyAsync(h.data(), sums.p, sizeof(i64) * ntiles, hipMemcpyDeviceToHost, s));
    PE_HIP(hipStreamSynchronize(s));
    i64 run = 0;
    for (i64 t = 0; t < ntiles; ++t) {
        i64 x = h[(size_t)t];
        h[(size_t)t] = run;
        run += x;
    }
    PE_HIP(hipMemcpyAsync(sums.p, h.data(), sizeof(i64) * ntiles, hipMemcpyHostToDevice, s));
    k_tile_scan<<<(unsigned)ntiles, SCAN_THREADS, 0, s>>>(d_data, n, sums.p);
    PE_LAUNCH_CHECK();
    PE_HIP(hipStreamSynchronize(s));  // sums and h leave scope
    return run;
}

}  // namespace pe
