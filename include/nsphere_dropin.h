/*
 * nsphere_dropin.h -- the reference's own hot-path C signatures, served by
 * the MI355X library.
 *
 * A C driver written against kris-sigurdson/NSphere v0.1.2 (src/c/nsphere.c)
 * drops its definitions of these functions and links libnsphere_gpu.so.
 * Caller-owned particles[5][npts] host arrays (r, v_r, ell, id as double, mu)
 * are copied to the GPU, processed and copied back on every call, exactly
 * where the reference would have processed them in place.  For a whole run
 * keep the state on the GPU with nsphere_gpu.h (ns_gpu_run) instead.
 *
 * Failure handling follows the reference's CLEAN_EXIT (nsphere.c:289-298):
 * the message is printed to stderr and the process exits with status 1.
 * The device is GPU 0 unless NSPHERE_GPU_DEVICE names another.
 */
#ifndef NSPHERE_DROPIN_H
#define NSPHERE_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* nsphere.c:880 (declaration), 8662-8665: sort every column by radius, in
 * place -- all five, mu (particles[4]) included, as nsphere.c:8613-8647 does.
 * Exact and stable (ties keep their input order, as with quadsort). */
void sort_particles(double **particles, int npts);

/* nsphere.c:903 (declaration), 8605-8652.  Every sortAlg value -- "quadsort",
 * "quadsort_parallel", "insertion", "insertion_parallel" or NULL -- gets the
 * exact stable sort; the reference's chunked variants approximate it. */
void sort_particles_with_alg(double **particles, int npts, const char *sortAlg);

/* The method-5 loop body, nsphere.c:5775-5838, which the reference inlines in
 * main(): sort (all five columns), inverse map (inverse_map may be NULL),
 * then every particle advanced by dt at its frozen rank with
 * doAdaptiveFullLeap or doAdaptiveFullLeviCivita and the constants of
 * nsphere.c:364-373, 5777-5782 (G_CONST, HALO_MASS, VEL_CONV_SQ, tolerances
 * 1e-7, max_subdiv 2^28, Richardson output, N_taumin 10, alpha 0.05).  No
 * work cap: a particle refines until it converges or reaches max_subdiv, as
 * in the reference.  The environment variable NSPHERE_GPU_BUDGET (force
 * evaluations per particle-step) sets the library's cap instead; each step
 * that stops particles at it prints their count on stderr. */
void nsphere_method5_step(double **particles, int npts, double dt, int *inverse_map);

/* Releases the GPU context the calls above keep between calls. */
void nsphere_dropin_finalize(void);

#ifdef __cplusplus
}
#endif
#endif
