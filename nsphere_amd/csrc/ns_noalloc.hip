// ns_noalloc.hip -- the allocation-free radial sort (ns_gpu_sort_particles_noalloc).
//
// The same stable radix sort and gather as ns_gpu_sort_particles, on arrays
// the caller owns, with the sort's workspace (key and payload ping-pong
// buffers, the permutation, digit totals and tile counts) carved out of one
// file-scope device array instead of a context's allocations.  Each GPU holds
// its own copy of the array from the library's load until the process ends;
// its device address is looked up once per device (hipGetSymbolAddress) and
// only that address is handed to kernels.  The reduce-then-scan sort needs no
// zeroed memory: every tile count and digit total is written before it is
// read, so the array carries nothing from one call to the next that matters.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ns_device.h"
#include "ns_internal.h"
#include "ns_kernels.h"

namespace ns {
namespace {

constexpr int64_t kMaxN = NS_SORT_NOALLOC_MAX;

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace layout for kMaxN particles.
constexpr size_t kKeysBytes = align256((size_t)kMaxN * sizeof(uint64_t));
constexpr size_t kValsBytes = align256((size_t)kMaxN * sizeof(uint32_t));
constexpr size_t kHistBytes = align256((size_t)2 * kKeyPasses * kRadix * sizeof(uint32_t));
constexpr size_t kTileBytes = align256((size_t)((kMaxN + kSortTile - 1) / kSortTile) * kRadix * sizeof(uint64_t));
constexpr size_t kCounterBytes = align256(64 * sizeof(uint32_t));
constexpr size_t kOffKeys0 = 0;
constexpr size_t kOffKeys1 = kOffKeys0 + kKeysBytes;
constexpr size_t kOffVals0 = kOffKeys1 + kKeysBytes;
constexpr size_t kOffVals1 = kOffVals0 + kValsBytes;
constexpr size_t kOffPerm = kOffVals1 + kValsBytes;
constexpr size_t kOffHist = kOffPerm + kValsBytes;
constexpr size_t kOffTiles = kOffHist + kHistBytes;
constexpr size_t kOffCounters = kOffTiles + kTileBytes;
constexpr size_t kScratchBytes = kOffCounters + kCounterBytes;
static_assert(kScratchBytes < ((size_t)32 << 20), "the allocation-free sort's scratch stays below 32 MiB");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The scratch array's address on `device` (current), looked up on first use.
int scratch_on(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NS_ERR_HIP, "ns_gpu_sort_particles_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                             device, hipGetErrorString(e));
        g_scratch[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[device];
    return NS_OK;
}

bool overlaps(const void *a, size_t an, const void *b, size_t bn)
{
    const char *x = static_cast<const char *>(a), *y = static_cast<const char *>(b);
    return x < y + bn && y < x + an;
}

} // namespace
} // namespace ns

extern "C" int ns_gpu_sort_particles_noalloc(const double *r, const double *v, const double *ell, const uint32_t *id,
                                             double *r_out, double *v_out, double *ell_out, uint32_t *id_out,
                                             int64_t n, void *stream)
{
    using namespace ns;
    clear_error();
    if (n < 0 || n > kMaxN)
        return set_error(NS_ERR_ARG, "ns_gpu_sort_particles_noalloc: n = %lld outside [0, %lld] (NS_SORT_NOALLOC_MAX)",
                         (long long)n, (long long)kMaxN);
    if (n == 0)
        return NS_OK;
    if (!r || !v || !ell || !id || !r_out || !v_out || !ell_out || !id_out)
        return set_error(NS_ERR_ARG, "ns_gpu_sort_particles_noalloc: null array");
    const size_t nd = (size_t)n * sizeof(double), ni = (size_t)n * sizeof(uint32_t);
    const void *ins[4] = {r, v, ell, id};
    const size_t insz[4] = {nd, nd, nd, ni};
    const void *outs[4] = {r_out, v_out, ell_out, id_out};
    const size_t outsz[4] = {nd, nd, nd, ni};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            if (overlaps(outs[a], outsz[a], ins[b], insz[b]) || (a < b && overlaps(outs[a], outsz[a], outs[b], outsz[b])))
                return set_error(NS_ERR_ARG, "ns_gpu_sort_particles_noalloc: output arrays overlap other arrays");
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices)
        return set_error(NS_ERR_NODEV, "ns_gpu_sort_particles_noalloc: no current HIP device");
    unsigned char *base = nullptr;
    if (int rc = scratch_on(device, &base))
        return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    SortWork w{};
    w.keys[0] = reinterpret_cast<uint64_t *>(base + kOffKeys0);
    w.keys[1] = reinterpret_cast<uint64_t *>(base + kOffKeys1);
    w.vals[0] = reinterpret_cast<uint32_t *>(base + kOffVals0);
    w.vals[1] = reinterpret_cast<uint32_t *>(base + kOffVals1);
    w.hist = reinterpret_cast<uint32_t *>(base + kOffHist);
    w.lookback = reinterpret_cast<uint64_t *>(base + kOffTiles);
    w.counters = reinterpret_cast<uint32_t *>(base + kOffCounters);
    w.epoch = 1;
    uint32_t *perm = reinterpret_cast<uint32_t *>(base + kOffPerm);
    hipError_t e = launch_radius_keys(r, w.keys[0], n, s);
    if (e == hipSuccess)
        e = radix_sort_perm(w, n, 64, perm, s, nullptr);
    if (e == hipSuccess)
        e = launch_gather_state<double>(r, v, ell, id, perm, r_out, v_out, ell_out, id_out, n, s);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_sort_particles_noalloc: %s", hipGetErrorString(e));
    return NS_OK;
}
