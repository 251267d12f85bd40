/*
 * ns_oracle.h -- CPU restatement of NSphere's per-step evolution loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (nsphere_amd) never links or calls it.
 *
 * Every function restates one reference function literally (same operation
 * order, same quirks); the cited file:line is in src/c/nsphere.c of
 * kris-sigurdson/NSphere v0.1.2.  Built twice from the same source:
 *   libns_oracle.so       -O2 -ffp-contract=off  (strict IEEE: the parity checker)
 *   libns_oracle_fast.so  -Ofast -march=native -ffp-contract=fast (the flags of
 *                         Makefile.Linux:63-74: the timed CPU baseline)
 */
#ifndef NS_ORACLE_H
#define NS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Loop parameters.  Field meaning follows nsphere.c:5777-5782 (tolerances,
 * max_subdiv, out_type, N_taumin, alpha) and the constants baked into the
 * force, nsphere.c:364-373 and 513 (vel_conv_sq * G * halo_mass). */
typedef struct {
    double dt;
    double G;            /* "grav" argument (G_CONST in the reference)          */
    double halo_mass;    /* HALO_MASS                                           */
    double vel_conv_sq;  /* VEL_CONV_SQ                                         */
    double radius_tol;
    double velocity_tol;
    double alpha;        /* alpha_param of the r_crit switch                    */
    int32_t max_subdiv;
    int32_t out_type;    /* 0 coarse, 1 fine, 2 "Richardson" 4F-3C              */
    int32_t N_taumin;
    int32_t identity_gravity; /* use_identity_gravity, nsphere.c:476           */
    /* Not in the reference: force evaluations allowed per particle and step
     * (0 = unlimited, the reference's behaviour).  Counted as the GPU library
     * counts them -- 2 for the first coarse leapfrog pass, 2N+1 per level,
     * 2N+2 / 4N+2 for the Levi-Civita coarse / fine passes -- and checked at
     * the same points, so capped results stay comparable bit for bit. */
    int64_t budget;
    /* Not in the reference (all its particles weigh M/npts): per-particle
     * masses' enclosed mass by rank, menc[i] below rank i, menc[npts + 1] the
     * rank-0 particle's mass; NULL = the reference's (rank/npts) M.  Used by
     * the force (M_enc in place of (rank/npts) M) and by the r_crit switch
     * (rank 0: a tenth of its own mass). */
    const double *menc;
} orc_params;

/* Per-particle refinement record written by the step (one byte):
 *   bits 0-5  deepest level K reached (N = 2^K), saturating at 63
 *   bit  6    the adaptive loop did not converge within max_subdiv
 *   bit  7    the particle took the Levi-Civita branch                         */
#define ORC_LVL_MASK 0x3F
#define ORC_LVL_NOCONV 0x40
#define ORC_LVL_LC 0x80

/* force helpers, nsphere.c:502-515, 532-535, 563-575, 596-599 */
double orc_gravitational_force(double r, int rank, int npts, const orc_params *p);
double orc_effective_angular_force(double r, double ell);
double orc_gravitational_force_rho_v(double rho, int rank, int npts, const orc_params *p);
double orc_effective_angular_force_rho_v(double rho, double ell);

/* nsphere.c:1550-1607 */
void orc_micro_leapfrog(int i, int npts, double r_in, double v_in, double ell,
                        double h, int N, int subSteps, const orc_params *p,
                        double *r_out, double *v_out);
/* nsphere.c:1648-1735; returns the level record (bits as above) */
int orc_adaptive_full_leap(int i, int npts, double r_in, double v_in, double ell,
                           double h, const orc_params *p, double *r_out, double *v_out);
/* nsphere.c:1990-2036 */
void orc_micro_lc(int i, int npts, double rho_in, double v_in, double t_in,
                  int64_t subSteps, double h_tau, const orc_params *p, double ell,
                  double *rho_out, double *v_out, double *t_out);
/* nsphere.c:2077-2186; returns the level record */
int orc_single_tau_step_lc(int i, int npts, double rho_in, double v_in, double t_in,
                           double h_guess, const orc_params *p, double ell,
                           double *rho_out, double *v_out, double *t_out);
/* nsphere.c:2231-2334; returns the level record (max over tau steps | LC bit);
 * *n_tau_steps (optional) receives the number of tau steps taken */
int orc_adaptive_full_lc(int i, int npts, double r_in, double v_in, double ell,
                         double dt, const orc_params *p, double *r_out, double *v_out,
                         int64_t *n_tau_steps);

/* IEEE total-order sort key: -0 folded onto +0, every NaN onto one key after +inf. */
uint64_t orc_radius_key(double r);

/* Stable sort of indices by radius (compare_particles, nsphere.c:8297-8306,
 * under the stable quadsort of nsphere.c:8356-8359).  perm[k] = input index of
 * the k-th smallest.  tiebreak_id != 0 breaks exact ties by id[] instead of by
 * input position. */
void orc_sort_perm(const double *r, const uint32_t *id, int64_t n, int tiebreak_id,
                   uint32_t *perm);
/* The reference default --sort 1 (nsphere.c:8475-8542): 24 chunks + seam
 * windows; approximate.  Used only to time the CPU baseline. */
void orc_sort_perm_chunked(const double *r, int64_t n, uint32_t *perm);

/* One method-5 step, nsphere.c:5775-5838, on SoA state in place:
 * sort (stable on r, or chunked when chunked_sort != 0), inverse map
 * (nsphere.c:5789-5794; inverse_map may be NULL), then the per-particle
 * adaptive integration at frozen rank.  level (may be NULL) receives the
 * per-particle level record in post-sort order. */
void orc_step(double *r, double *v, double *ell, uint32_t *id, int64_t npts,
              const orc_params *p, int chunked_sort, int tiebreak_id,
              int32_t *inverse_map, uint8_t *level);

/* Integration only (no sort): particles are taken to be in rank order already. */
void orc_integrate(double *r, double *v, const double *ell, int64_t npts,
                   const orc_params *p, uint8_t *level);

/* Shell-model energy with the rank-based enclosed mass of nsphere.c:513:
 * K = sum m (v^2 + ell^2/r^2)/2,  W = -sum (vel_conv_sq G) (i/npts) M m / r,
 * m = M/npts, i = rank.  Summed in long double. */
/* Enclosed mass for per-particle masses given in rank order: out[i] = the
 * sum of mass[0..i) to about 2^-64 relative (compensated long double),
 * rounded once to double; out[npts] = the total, out[npts + 1] = mass[0] --
 * the layout of orc_params.menc. */
void orc_mass_prefix(const double *mass, int64_t npts, double *out);
/* orc_energy with per-particle masses (rank order; NULL = equal masses). */
void orc_energy_masses(const double *r, const double *v, const double *ell, const double *mass,
                       int64_t npts, const orc_params *p, double *K, double *W);
void orc_energy(const double *r, const double *v, const double *ell, int64_t npts,
                const orc_params *p, double *K, double *W);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
