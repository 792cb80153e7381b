// Engine context, status mapping, combinatorics, host index, SpookyHash and
// the integral container (FCIDUMP I/O).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <mutex>
#include <regex>
#include <sstream>

#include "host.hpp"

namespace ph {

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
[[noreturn]] void raise_status(int rc) {
    std::string msg = pyci_last_error();
    switch (rc) {
        case PYCI_EINVAL:
        case PYCI_EUNSUPPORTED: throw py::value_error(msg);
        case PYCI_ETYPE: throw py::type_error(msg);
        case PYCI_ENOMEM: throw std::bad_alloc();
        case PYCI_ENOCONV: throw std::runtime_error(msg.empty() ? "did not converge" : msg);
        default: throw std::runtime_error("pyci_amd engine: " + msg);
    }
}

static std::mutex g_ctx_mutex;
static pyci_ctx *g_ctx = nullptr;

pyci_ctx *context() {
    std::lock_guard<std::mutex> lk(g_ctx_mutex);
    if (!g_ctx) {
        int dev = 0;
        if (const char *e = std::getenv("PYCI_AMD_DEVICE")) dev = std::atoi(e);
        pyci_ctx *c = nullptr;
        check(pyci_ctx_create(dev, &c));
        g_ctx = c;
    }
    return g_ctx;
}

// ---------------------------------------------------------------------------
// Combinatorics and bits (common.cpp:41-153 semantics)
// ---------------------------------------------------------------------------
i64 binomial(i64 n, i64 k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    long double approx = 1.0L;
    for (i64 d = 1; d <= k; ++d) approx = approx * (long double)(n - k + d) / (long double)d;
    if (approx > 9.2e18L) return std::numeric_limits<i64>::max();
    unsigned __int128 b = 1;
    for (i64 d = 1; d <= k; ++d) b = b * (unsigned __int128)(n - k + d) / (unsigned __int128)d;
    return (i64)b;
}

void fill_hartreefock_det(i64 nocc, u64 *det) {
    i64 i = 0;
    while (nocc >= 64) {
        det[i++] = ~0ULL;
        nocc -= 64;
    }
    if (nocc) det[i] = (1ULL << nocc) - 1ULL;
}

void fill_det(i64 nocc, const i64 *occs, u64 *det) {
    for (i64 i = 0; i < nocc; ++i) det[occs[i] / 64] |= 1ULL << (occs[i] % 64);
}

void fill_occs(i64 nword, const u64 *det, i64 *occs) {
    i64 j = 0;
    for (i64 w = 0; w < nword; ++w)
        for (u64 x = det[w]; x; x &= x - 1) occs[j++] = __builtin_ctzll(x) + 64 * w;
}

void fill_virs(i64 nword, i64 nbasis, const u64 *det, i64 *virs) {
    i64 j = 0;
    for (i64 w = 0; w < nword; ++w) {
        i64 bits = std::min<i64>(64, nbasis - 64 * w);
        u64 mask = bits >= 64 ? ~0ULL : ((1ULL << bits) - 1ULL);
        for (u64 x = det[w] ^ mask; x; x &= x - 1) virs[j++] = __builtin_ctzll(x) + 64 * w;
    }
}

void next_colex(i64 *idx) {
    i64 i = 0;
    while (idx[i + 1] - idx[i] == 1) {
        idx[i] = i;
        ++i;
    }
    ++idx[i];
}

i64 rank_colex(i64 nbasis, i64 nocc, const u64 *det) {
    i64 r = 0, k = 0;
    for (i64 p = 0; p < nbasis && k < nocc; ++p)
        if (det[p / 64] >> (p % 64) & 1ULL) r += binomial(p, ++k);
    return r;
}

// ---------------------------------------------------------------------------
// Host index
// ---------------------------------------------------------------------------
u64 HostIndex::hash(const u64 *d, i64 nw) {
    u64 h = 0x9e3779b97f4a7c15ULL;
    for (i64 i = 0; i < nw; ++i) {
        u64 x = d[i] + 0x632be59bd9b4e019ULL * (u64)(i + 1);
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdULL;
        x ^= x >> 33;
        h = (h ^ x) * 0xc4ceb9fe1a85ec53ULL;
        h ^= h >> 29;
    }
    return h;
}

i64 HostIndex::find(const std::vector<u64> &dets, i64 nw, const u64 *d) const {
    if (slots.empty()) return -1;
    const u64 m = slots.size() - 1;
    u64 h = hash(d, nw) & m;
    for (;;) {
        i64 v = slots[h];
        if (v < 0) return -1;
        if (std::memcmp(&dets[(size_t)(v * nw)], d, sizeof(u64) * nw) == 0) return v;
        h = (h + 1) & m;
    }
}

void HostIndex::insert(const std::vector<u64> &dets, i64 nw, i64 idx) {
    if (2 * (count + 1) > (i64)slots.size()) {
        size_t cap = std::max<size_t>(64, slots.size() * 2);
        while ((i64)cap < 2 * (count + 1)) cap *= 2;
        slots.assign(cap, -1);
        i64 c = count;
        count = 0;
        for (i64 v = 0; v < c; ++v) insert(dets, nw, v);
    }
    const u64 m = slots.size() - 1;
    u64 h = hash(&dets[(size_t)(idx * nw)], nw) & m;
    while (slots[h] >= 0) h = (h + 1) & m;
    slots[h] = idx;
    ++count;
}

void HostIndex::rebuild(const std::vector<u64> &dets, i64 nw, i64 n) {
    size_t cap = 64;
    while ((i64)cap < 2 * n + 2) cap *= 2;
    slots.assign(cap, -1);
    count = 0;
    const u64 m = cap - 1;
    for (i64 v = 0; v < n; ++v) {
        u64 h = hash(&dets[(size_t)(v * nw)], nw) & m;
        while (slots[h] >= 0) h = (h + 1) & m;
        slots[h] = v;
    }
    count = n;
}

// ---------------------------------------------------------------------------
// SpookyHash V2 (Bob Jenkins, public domain), 128-bit variant.
// ---------------------------------------------------------------------------
namespace {
constexpr u64 SC = 0xdeadbeefdeadbeefULL;
inline u64 rot(u64 x, int k) { return (x << k) | (x >> (64 - k)); }

inline void short_mix(u64 &h0, u64 &h1, u64 &h2, u64 &h3) {
    h2 = rot(h2, 50), h2 += h3, h0 ^= h2;
    h3 = rot(h3, 52), h3 += h0, h1 ^= h3;
    h0 = rot(h0, 30), h0 += h1, h2 ^= h0;
    h1 = rot(h1, 41), h1 += h2, h3 ^= h1;
    h2 = rot(h2, 54), h2 += h3, h0 ^= h2;
    h3 = rot(h3, 48), h3 += h0, h1 ^= h3;
    h0 = rot(h0, 38), h0 += h1, h2 ^= h0;
    h1 = rot(h1, 37), h1 += h2, h3 ^= h1;
    h2 = rot(h2, 62), h2 += h3, h0 ^= h2;
    h3 = rot(h3, 34), h3 += h0, h1 ^= h3;
    h0 = rot(h0, 5), h0 += h1, h2 ^= h0;
    h1 = rot(h1, 36), h1 += h2, h3 ^= h1;
}

inline void short_end(u64 &h0, u64 &h1, u64 &h2, u64 &h3) {
    h3 ^= h2, h2 = rot(h2, 15), h3 += h2;
    h0 ^= h3, h3 = rot(h3, 52), h0 += h3;
    h1 ^= h0, h0 = rot(h0, 26), h1 += h0;
    h2 ^= h1, h1 = rot(h1, 51), h2 += h1;
    h3 ^= h2, h2 = rot(h2, 28), h3 += h2;
    h0 ^= h3, h3 = rot(h3, 9), h0 += h3;
    h1 ^= h0, h0 = rot(h0, 47), h1 += h0;
    h2 ^= h1, h1 = rot(h1, 54), h2 += h1;
    h3 ^= h2, h2 = rot(h2, 32), h3 += h2;
    h0 ^= h3, h3 = rot(h3, 25), h0 += h3;
    h1 ^= h0, h0 = rot(h0, 63), h1 += h0;
}

inline u64 ld64(const uint8_t *p) {
    u64 x;
    std::memcpy(&x, p, 8);
    return x;
}

inline uint32_t ld32(const uint8_t *p) {
    uint32_t x;
    std::memcpy(&x, p, 4);
    return x;
}

void spooky_short(const uint8_t *p, size_t len, u64 &h1, u64 &h2) {
    size_t rem = len % 32;
    u64 a = h1, b = h2, c = SC, d = SC;
    if (len > 15) {
        const uint8_t *end = p + (len / 32) * 32;
        for (; p < end; p += 32) {
            c += ld64(p), d += ld64(p + 8);
            short_mix(a, b, c, d);
            a += ld64(p + 16), b += ld64(p + 24);
        }
        if (rem >= 16) {
            c += ld64(p), d += ld64(p + 8);
            short_mix(a, b, c, d);
            p += 16;
            rem -= 16;
        }
    }
    d += ((u64)len) << 56;
    switch (rem) {
        case 15: d += ((u64)p[14]) << 48; [[fallthrough]];
        case 14: d += ((u64)p[13]) << 40; [[fallthrough]];
        case 13: d += ((u64)p[12]) << 32; [[fallthrough]];
        case 12: d += ld32(p + 8); c += ld64(p); break;
        case 11: d += ((u64)p[10]) << 16; [[fallthrough]];
        case 10: d += ((u64)p[9]) << 8; [[fallthrough]];
        case 9: d += (u64)p[8]; [[fallthrough]];
        case 8: c += ld64(p); break;
        case 7: c += ((u64)p[6]) << 48; [[fallthrough]];
        case 6: c += ((u64)p[5]) << 40; [[fallthrough]];
        case 5: c += ((u64)p[4]) << 32; [[fallthrough]];
        case 4: c += ld32(p); break;
        case 3: c += ((u64)p[2]) << 16; [[fallthrough]];
        case 2: c += ((u64)p[1]) << 8; [[fallthrough]];
        case 1: c += (u64)p[0]; break;
        case 0: c += SC, d += SC;
    }
    short_end(a, b, c, d);
    h1 = a, h2 = b;
}

inline void mix(const u64 *x, u64 *s) {
    static const int r[12] = {11, 32, 43, 31, 17, 28, 39, 57, 55, 54, 22, 46};
    for (int i = 0; i < 12; ++i) {
        s[i] += x[i];
        s[(i + 2) % 12] ^= s[(i + 10) % 12];
        s[(i + 11) % 12] ^= s[i];
        s[i] = rot(s[i], r[i]);
        s[(i + 11) % 12] += s[(i + 1) % 12];
    }
}

inline void end_partial(u64 *h) {
    static const int r[12] = {44, 15, 34, 21, 38, 33, 10, 13, 38, 53, 42, 54};
    for (int i = 0; i < 12; ++i) {
        h[(i + 11) % 12] += h[(i + 1) % 12];
        h[(i + 2) % 12] ^= h[(i + 11) % 12];
        h[(i + 1) % 12] = rot(h[(i + 1) % 12], r[i]);
    }
}
}  // namespace

std::pair<u64, u64> spooky128(const void *msg, size_t len, u64 seed1, u64 seed2) {
    const uint8_t *p = (const uint8_t *)msg;
    u64 h1 = seed1, h2 = seed2;
    if (len < 192) {
        spooky_short(p, len, h1, h2);
        return {h1, h2};
    }
    u64 s[12];
    for (int i = 0; i < 12; i += 3) s[i] = seed1, s[i + 1] = seed2, s[i + 2] = SC;
    const size_t nblk = len / 96;
    u64 buf[12];
    for (size_t b = 0; b < nblk; ++b) {
        std::memcpy(buf, p + b * 96, 96);
        mix(buf, s);
    }
    const size_t done = nblk * 96, rem = len - done;
    uint8_t last[96];
    std::memcpy(last, p + done, rem);
    std::memset(last + rem, 0, 96 - rem);
    last[95] = (uint8_t)rem;
    std::memcpy(buf, last, 96);
    for (int i = 0; i < 12; ++i) s[i] += buf[i];
    end_partial(s);
    end_partial(s);
    end_partial(s);
    return {s[0], s[1]};
}

// ---------------------------------------------------------------------------
// Integrals
// ---------------------------------------------------------------------------
void SQuantOp::derive() {
    const i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1;
    h = Array<double>(n1);
    v = Array<double>({n1, n1});
    w = Array<double>({n1, n1});
    const double *g = two_mo.data();
    const double *o = one_mo.data();
    double *ph_ = h.mutable_data(), *pv = v.mutable_data(), *pw = w.mutable_data();
    for (i64 i = 0; i < n1; ++i) {
        ph_[i] = o[i * (n1 + 1)];
        for (i64 j = 0; j < n1; ++j) {
            pv[i * n1 + j] = g[i * n3 + i * n2 + j * n1 + j];
            pw[i * n1 + j] = g[i * n3 + j * n2 + i * n1 + j] * 2 - g[i * n3 + j * n2 + j * n1 + i];
        }
    }
}

SQuantOp::SQuantOp(double e, Array<double> mo1, Array<double> mo2) : ecore(e), one_mo(mo1), two_mo(mo2) {
    if (one_mo.ndim() != 2 || one_mo.shape(0) != one_mo.shape(1))
        throw std::invalid_argument("one_mo must be a square (n, n) array");
    nbasis = one_mo.shape(0);
    if (two_mo.ndim() != 4 || two_mo.shape(0) != nbasis || two_mo.shape(1) != nbasis ||
        two_mo.shape(2) != nbasis || two_mo.shape(3) != nbasis)
        throw std::invalid_argument("two_mo must have shape (n, n, n, n)");
    derive();
}

// FCIDUMP reader (squantop.cpp:79-161 semantics): header lines up to the one
// holding "&END" or "/", NORB required, UHF rejected; then "x i j k l" records.
SQuantOp::SQuantOp(const std::string &filename) {
    std::ifstream f(filename);
    if (f.fail()) throw std::ios_base::failure("Failed to read the FCIDUMP file " + filename);
    std::string header, line;
    bool ended = false;
    while (std::getline(f, line)) {
        if (line.find("&END") != std::string::npos || line.find('/') != std::string::npos) {
            ended = true;
            break;
        }
        header += " " + line;
    }
    if (!ended) throw std::ios_base::failure("FCIDUMP has the wrong header");
    std::smatch m;
    if (!std::regex_search(header, m, std::regex(R"(NORB[ ]*=[ ]*(\d+))")))
        throw std::invalid_argument("NORB is not found.");
    nbasis = std::stol(m[1]);
    if (std::regex_search(header, m, std::regex(R"(UHF[ ]*=[ .]*(FALSE|TRUE))", std::regex::icase))) {
        std::string u = m[1];
        std::transform(u.begin(), u.end(), u.begin(), ::toupper);
        if (u == "TRUE") throw std::runtime_error("Unrestricted FCIDUMP not implemented");
    }
    const i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1;
    one_mo = Array<double>({n1, n1});
    two_mo = Array<double>({n1, n1, n1, n1});
    double *o = one_mo.mutable_data(), *g = two_mo.mutable_data();
    std::fill(o, o + n2, 0.0);
    std::fill(g, g + n3 * n1, 0.0);
    ecore = 0.0;
    double x;
    i64 i, j, k, l;
    while (f >> x >> i >> j >> k >> l) {
        if (i && j && k && l) {
            --i, --j, --k, --l;
            if (i < 0 || j < 0 || k < 0 || l < 0 || i >= n1 || j >= n1 || k >= n1 || l >= n1)
                throw std::invalid_argument("FCIDUMP index out of range");
            const i64 idx[8][4] = {{i, k, j, l}, {k, i, l, j}, {j, k, i, l}, {i, l, j, k},
                                   {j, l, i, k}, {l, j, k, i}, {k, j, l, i}, {l, i, k, j}};
            for (auto &t : idx) g[t[0] * n3 + t[1] * n2 + t[2] * n1 + t[3]] = x;
        } else if (i && j) {
            --i, --j;
            if (i < 0 || j < 0 || i >= n1 || j >= n1) throw std::invalid_argument("FCIDUMP index out of range");
            o[i * n1 + j] = x;
            o[j * n1 + i] = x;
        } else {
            ecore = x;
        }
    }
    derive();
}

// FCIDUMP writer (squantop.cpp:185-223 format).
void SQuantOp::to_file(const std::string &filename, i64 nelec, i64 ms2, double tol) const {
    std::ofstream f(filename);
    if (f.fail()) throw std::ios_base::failure("Failed to open the FCIDUMP file " + filename);
    const i64 n1 = nbasis, n2 = n1 * n1, n3 = n2 * n1;
    f << "&FCIDUMP\nNORB=" << nbasis << ",\nNELEC=" << nelec << ",\nMS2=" << ms2 << ",\nUHF=.FALSE.,\nORBSYM=";
    for (i64 i = 0; i < nbasis; ++i) f << "1,";
    f << "\nISYM=1,\n&END\n";
    const double *g = two_mo.data(), *o = one_mo.data();
    auto put = [&](double val, i64 a, i64 b, i64 c, i64 d) {
        f << std::setw(28) << std::setprecision(20) << std::scientific << val << ' ' << a << ' ' << b << ' ' << c
          << ' ' << d << "\n";
    };
    for (i64 i = 0; i < n1; ++i)
        for (i64 j = 0; j <= i; ++j)
            for (i64 k = 0; k < n1; ++k)
                for (i64 l = 0; l <= k; ++l)
                    if ((i * (i + 1)) / 2 + j >= (k * (k + 1)) / 2 + l) {
                        double val = g[i * n3 + k * n2 + j * n1 + l];
                        if (std::abs(val) > tol) put(val, i + 1, j + 1, k + 1, l + 1);
                    }
    for (i64 i = 0; i < n1; ++i)
        for (i64 j = 0; j <= i; ++j) {
            double val = o[i * n1 + j];
            if (std::abs(val) > tol) put(val, i + 1, j + 1, 0, 0);
        }
    put(ecore, 0, 0, 0, 0);
}

pyci_ham *SQuantOp::device() {
    if (!dev) {
        pyci_ctx *ctx = context();
        pyci_ham *hp = nullptr;
        check(pyci_ham_create(ctx, nbasis, ecore, one_mo.data(), two_mo.data(), &hp));
        dev = std::shared_ptr<pyci_ham>(hp, pyci_ham_destroy);
    }
    return dev.get();
}

}  // namespace ph
