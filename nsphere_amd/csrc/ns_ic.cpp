// ns_ic.cpp -- seeded initial conditions for the evolution loop (host side).
//
// The reference samples its initial conditions from a GSL spline of its own
// cored profile, seeded from time+pid (nsphere.c:4656-4955, 4699-4700).  The
// workloads here are instead closed-form Plummer and Hernquist spheres
// (ns_ic_sampler.h), drawn on the host with OpenMP here or on the device by
// ns_gpu_generate_ic.

#include <cstdint>
#include <cstdio>

#include "nsphere_gpu.h"
#include "ns_ic_sampler.h"
#include "ns_internal.h"

extern "C" void ns_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    ns::ic::philox4x32_10(ctr, key, out);
}

extern "C" int ns_ic_generate(int kind, int64_t n, uint64_t seed, double rmax, double lowl_frac,
                              double lowl_scale, double *r, double *v, double *ell, uint32_t *id)
{
    if (n < 0 || n > (int64_t)UINT32_MAX || kind < 0 || kind > 2 || !(rmax > 0.0))
        return ns::set_error(NS_ERR_ARG, "ns_ic_generate: bad kind/n/rmax");
    if (n > 0 && (!r || !v || !ell || !id))
        return ns::set_error(NS_ERR_ARG, "ns_ic_generate: null output array");
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        ns::ic::sample_particle(kind, seed, (uint64_t)i, rmax, lowl_frac, lowl_scale, r[i], v[i], ell[i]);
        id[i] = (uint32_t)i;
    }
    return NS_OK;
}

// write_initial_conditions, nsphere.c:2386-2416: npts, then the five values
// of each particle in turn.
extern "C" int ns_ic_file_write(const char *path, int32_t npts, const double *records)
{
    if (!path || npts < 0 || (npts > 0 && !records))
        return ns::set_error(NS_ERR_ARG, "ns_ic_file_write: bad argument");
    FILE *fp = fopen(path, "wb");
    if (!fp)
        return ns::set_error(NS_ERR_IO, "ns_ic_file_write: cannot open %s for writing", path);
    bool ok = fwrite(&npts, sizeof(int32_t), 1, fp) == 1;
    if (ok && npts > 0)
        ok = fwrite(records, sizeof(double), (size_t)npts * 5, fp) == (size_t)npts * 5;
    ok = (fclose(fp) == 0) && ok;
    return ok ? NS_OK : ns::set_error(NS_ERR_IO, "ns_ic_file_write: short write to %s", path);
}

// read_initial_conditions, nsphere.c:2445-2487.
extern "C" int ns_ic_file_read(const char *path, int32_t *npts, double *records)
{
    if (!path || !npts)
        return ns::set_error(NS_ERR_ARG, "ns_ic_file_read: bad argument");
    FILE *fp = fopen(path, "rb");
    if (!fp)
        return ns::set_error(NS_ERR_IO, "ns_ic_file_read: cannot open %s for reading", path);
    int32_t file_npts = 0;
    if (fread(&file_npts, sizeof(int32_t), 1, fp) != 1 || file_npts < 0) {
        fclose(fp);
        return ns::set_error(NS_ERR_IO, "ns_ic_file_read: failed to read npts from %s", path);
    }
    if (!records) {
        fclose(fp);
        *npts = file_npts;
        return NS_OK;
    }
    if (file_npts != *npts) {
        fclose(fp);
        return ns::set_error(NS_ERR_IO, "ns_ic_file_read: %s holds %d particles, %d expected", path, (int)file_npts,
                             (int)*npts);
    }
    const size_t want = (size_t)file_npts * 5;
    const size_t got = want ? fread(records, sizeof(double), want, fp) : 0;
    fclose(fp);
    if (got != want)
        return ns::set_error(NS_ERR_IO, "ns_ic_file_read: partial read at particle %zu of %s", got / 5, path);
    return NS_OK;
}
