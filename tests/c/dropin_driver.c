/*
 * dropin_driver.c -- a C host program written against the reference's own
 * function signatures (nsphere_dropin.h), linked to libnsphere_gpu.so.
 *
 *   dropin_driver <in.bin> <out.bin> <dt> <nsteps>
 *
 * in.bin / out.bin: int64 npts, then npts doubles each of r, v_r, ell, id,
 * mu (particles[0..4]).  Runs the reference's initial exact sort
 * (sort_particles_with_alg with "quadsort", nsphere.c:5339-5342), then nsteps
 * method-5 steps, and writes the five columns and the last inverse map (npts
 * int32) out.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "nsphere_dropin.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s in.bin out.bin dt nsteps\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f)
        return 3;
    int64_t n = 0;
    if (fread(&n, sizeof n, 1, f) != 1 || n <= 0 || n > (1 << 30))
        return 4;
    double *particles[5];
    for (int c = 0; c < 5; c++) {
        particles[c] = (double *)calloc((size_t)n, sizeof(double));
        if (!particles[c])
            return 5;
    }
    for (int c = 0; c < 5; c++)
        if (fread(particles[c], sizeof(double), (size_t)n, f) != (size_t)n)
            return 6;
    fclose(f);
    int *inverse_map = (int *)calloc((size_t)n, sizeof(int));
    const double dt = atof(argv[3]);
    const int nsteps = atoi(argv[4]);

    sort_particles_with_alg(particles, (int)n, "quadsort");
    for (int s = 0; s < nsteps; s++)
        nsphere_method5_step(particles, (int)n, dt, inverse_map);
    nsphere_dropin_finalize();

    f = fopen(argv[2], "wb");
    if (!f)
        return 7;
    fwrite(&n, sizeof n, 1, f);
    for (int c = 0; c < 5; c++)
        fwrite(particles[c], sizeof(double), (size_t)n, f);
    fwrite(inverse_map, sizeof(int), (size_t)n, f);
    fclose(f);
    printf("dropin_driver: %lld particles, %d steps\n", (long long)n, nsteps);
    return 0;
}
