/*
 * ref_quadsort_shim.c -- caller for the reference's own vendored sort.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
 * the quadsort 1.2.1.3 sources where they lie in the reference tree
 * (src/c/nsphere_sort.h -> quadsort.h -> quadsort.c); the output goes to
 * oracle/_ref/ and is never part of the product.  It sorts row pointers with
 * a radius comparator of the same semantics as compare_particles
 * (nsphere.c:8297-8306), exactly as sort_particles_with_alg(..., "quadsort")
 * does (nsphere.c:8605-8652, 8356-8359), and reports the permutation.
 */
#include <stdint.h>
#include <stdlib.h>
#include "nsphere_sort.h"

static int cmp_rows(const void *a, const void *b)
{
    const double *ra = *(double *const *)a;
    const double *rb = *(double *const *)b;
    if (ra[0] < rb[0])
        return -1;
    if (ra[0] > rb[0])
        return 1;
    return 0;
}

/* perm[k] = input index of the k-th row after quadsort (the reference's
 * exact, stable radial sort). Returns 0 on success. */
int ref_quadsort_perm(const double *r, int64_t n, uint32_t *perm)
{
    double *rows = (double *)malloc((size_t)(n > 0 ? n : 1) * 2 * sizeof(double));
    double **cols = (double **)malloc((size_t)(n > 0 ? n : 1) * sizeof(double *));
    if (!rows || !cols) {
        free(rows);
        free(cols);
        return 1;
    }
    for (int64_t i = 0; i < n; i++) {
        rows[2 * i] = r[i];
        rows[2 * i + 1] = (double)i;
        cols[i] = &rows[2 * i];
    }
    quadsort(cols, (size_t)n, sizeof(double *), cmp_rows);
    for (int64_t i = 0; i < n; i++)
        perm[i] = (uint32_t)cols[i][1];
    free(rows);
    free(cols);
    return 0;
}
