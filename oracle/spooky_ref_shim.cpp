// C entry point over the reference's SpookyHash V2 (compiled from the reference
// sources where they lie; see oracle/Makefile). Test infrastructure only: it
// checks the host-side rank_det implementation in pyci_amd.
#include <stdint.h>
#include <SpookyV2.h>

extern "C" void spooky_ref_hash128(const void *msg, size_t len, uint64_t *h1, uint64_t *h2) {
    uint64 a = *h1, b = *h2;
    SpookyHash::Hash128(msg, len, &a, &b);
    *h1 = a;
    *h2 = b;
}
