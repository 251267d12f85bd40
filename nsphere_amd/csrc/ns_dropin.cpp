// ns_dropin.cpp -- the reference's hot-path C signatures over the C-ABI.
//
// Host code only: every call goes through nsphere_gpu.h.  One context is kept
// per process (recreated when npts changes); failures print and exit(1) as
// the reference's CLEAN_EXIT does (nsphere.c:289-298).
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "nsphere_dropin.h"
#include "nsphere_gpu.h"

namespace {

std::mutex g_mu;
ns_gpu_ctx *g_ctx = nullptr;
long long g_npts = -1;

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
    if (inverse_map)
        check(ns_gpu_inverse_map(c, inverse_map), "ns_gpu_inverse_map");
    check(ns_gpu_download(c, particles), "ns_gpu_download");
}

extern "C" void nsphere_dropin_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx)
        ns_gpu_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_npts = -1;
}
