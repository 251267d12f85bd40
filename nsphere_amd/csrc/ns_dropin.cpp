// ns_dropin.cpp -- the reference's hot-path C signatures over the C-ABI.
//
// Host code only: every call goes through nsphere_gpu.h.  One context is kept
// per process (recreated when npts changes); failures print and exit(1) as
// the reference's CLEAN_EXIT does (nsphere.c:289-298).
//
// All five columns move together, as in sort_particles_with_alg
// (nsphere.c:8613-8647): r, v_r, ell and id through the device state, mu
// (particles[4], which the library does not keep) by ns_gpu_permute_column.
//
// The step runs with no work cap, so it stops refining only where the
// reference does (max_subdiv = 2^28 per level, nsphere.c:5779, and at most
// 1e8 tau steps).  NSPHERE_GPU_BUDGET=<force evaluations per particle-step>
// sets the library's cap instead; every step that stops particles at it then
// says so on stderr, since the reference has no such cap.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "nsphere_dropin.h"
#include "nsphere_gpu.h"

namespace {

std::mutex g_mu;
ns_gpu_ctx *g_ctx = nullptr;
long long g_npts = -1;
long long g_budget = 0;  // 0: no cap

[[noreturn]] void die(const char *what)
{
    std::fprintf(stderr, "nsphere (MI355X): %s failed: %s\n", what, ns_gpu_last_error());
    std::exit(1);
}

void check(int rc, const char *what)
{
    if (rc != NS_OK)
        die(what);
}

ns_gpu_ctx *context_for(int npts)
{
    if (g_ctx && g_npts == npts)
        return g_ctx;
    if (g_ctx) {
        ns_gpu_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    int device = 0;
    if (const char *s = std::getenv("NSPHERE_GPU_DEVICE"))
        device = std::atoi(s);
    check(ns_gpu_ctx_create(npts, device, &g_ctx), "ns_gpu_ctx_create");
    g_npts = npts;
    g_budget = 0;
    if (const char *s = std::getenv("NSPHERE_GPU_BUDGET"))
        g_budget = std::atoll(s);
    check(ns_gpu_set_option(g_ctx, NS_OPT_BUDGET, g_budget), "NS_OPT_BUDGET (NSPHERE_GPU_BUDGET)");
    return g_ctx;
}

} // namespace

extern "C" void sort_particles_with_alg(double **particles, int npts, const char *sortAlg)
{
    (void)sortAlg;
    std::lock_guard<std::mutex> lk(g_mu);
    if (npts <= 0)
        return;
    ns_gpu_ctx *c = context_for(npts);
    check(ns_gpu_upload(c, particles), "ns_gpu_upload");
    check(ns_gpu_sort_particles(c), "ns_gpu_sort_particles");
    check(ns_gpu_download(c, particles), "ns_gpu_download");
    check(ns_gpu_permute_column(c, particles[4], particles[4]), "ns_gpu_permute_column");
}

extern "C" void sort_particles(double **particles, int npts)
{
    sort_particles_with_alg(particles, npts, "quadsort_parallel");
}

extern "C" void nsphere_method5_step(double **particles, int npts, double dt, int *inverse_map)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (npts <= 0)
        return;
    ns_gpu_ctx *c = context_for(npts);
    ns_step_params p;
    ns_step_params_default(&p);
    p.dt = dt;
    check(ns_gpu_upload(c, particles), "ns_gpu_upload");
    check(ns_gpu_step(c, &p), "ns_gpu_step");
    check(ns_gpu_sync(c), "ns_gpu_step");
    if (g_budget > 0) {
        uint64_t hits = 0;
        check(ns_gpu_cap_hits(c, &hits, 1), "ns_gpu_cap_hits");
        if (hits > 0)
            std::fprintf(stderr,
                         "nsphere (MI355X): %" PRIu64 " particles stopped at NSPHERE_GPU_BUDGET=%lld force "
                         "evaluations this step (the reference has no such cap)\n",
                         hits, g_budget);
    }
    if (inverse_map)
        check(ns_gpu_inverse_map(c, inverse_map), "ns_gpu_inverse_map");
    check(ns_gpu_download(c, particles), "ns_gpu_download");
    check(ns_gpu_permute_column(c, particles[4], particles[4]), "ns_gpu_permute_column");
}

extern "C" void nsphere_dropin_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx)
        ns_gpu_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_npts = -1;
}
