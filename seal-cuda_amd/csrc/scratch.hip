// scratch.hip -- a fixed device scratch array for the entry point that allocates nothing
// (shc_relinearize_noalloc), for callers that cannot call hipMalloc: code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation.
//
// The array lives in the code object (no allocation call), one copy per device, held
// until the process ends.  Host code never passes the array's name to a kernel: in
// host code it is a stand-in, not a device address.  Its device address is looked up
// once per device with hipGetSymbolAddress (with that device current) and kept; the
// kernels receive that pointer, as they receive a caller's workspace elsewhere.
#include "scratch.h"
#include "rns_ops.h"
#include <mutex>

static __device__ __align__(256) unsigned char scratch[shc::NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};

}  // namespace

namespace shc {

void *noalloc_scratch_setup() {
    int dev = 0;
    SHC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) fail(ERR_OUT_OF_RANGE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        SHC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

void relinearize_noalloc(const Context &c, uint64_t *ct, const uint64_t *key, uint32_t count, uint32_t L,
                         hipStream_t st) {
    const size_t need = keyswitch_workspace(c, count, L);
    if (need > NOALLOC_SCRATCH_BYTES)
        fail(ERR_OUT_OF_RANGE, "relinearize_noalloc: the key switch needs " + std::to_string(need) +
                                   " bytes of scratch, more than the fixed " +
                                   std::to_string(NOALLOC_SCRATCH_BYTES) + " (use shc_relinearize with a workspace)");
    void *ws = noalloc_scratch_setup();
    const size_t poly = (size_t)L * c.n;
    switch_key(c, ct + 2 * poly, 3 * poly, ct, 3 * poly, key, count, L, ws, NOALLOC_SCRATCH_BYTES, st);
}

}  // namespace shc
