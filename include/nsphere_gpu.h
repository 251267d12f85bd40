/*
 * nsphere_gpu.h -- C-ABI of the MI355X-native NSphere per-step evolution loop.
 *
 * One context owns one GPU's copy of the particle state (SoA r, v_r, ell in
 * float64 plus a uint32 id, kept in radius-rank order as the reference keeps
 * particles[5][npts]) and one HIP stream.  A context is not thread-safe.
 *
 * Every entry point returns 0 on success or a positive NS_ERR_* code; the
 * message of the last failure on the calling thread is ns_gpu_last_error().
 * Nothing here terminates the process (the drop-in layer in nsphere_dropin.h
 * does, as the reference does on failure).
 *
 * Reference interfaces replaced (kris-sigurdson/NSphere v0.1.2, src/c/nsphere.c):
 *   ns_gpu_sort_particles  <- sort_particles / sort_particles_with_alg
 *                             (declared 880, 903; defined 8662-8665, 8605-8652)
 *   ns_gpu_inverse_map     <- the inverse-map loop, 5789-5794
 *   ns_gpu_integrate       <- the per-particle loop, 5796-5838, which calls
 *                             doAdaptiveFullLeap (1648-1735) and
 *                             doAdaptiveFullLeviCivita (2231-2334)
 *   ns_gpu_step            <- the whole method-5 body, 5775-5838
 *   ns_gpu_upload/download <- the caller-owned double *particles[5] layout,
 *                             4656-4682 (r, v_r, ell, id, mu; mu is not used)
 */
#ifndef NSPHERE_GPU_H
#define NSPHERE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NS_OK = 0,
    NS_ERR_ARG = 1,     /* invalid argument (null pointer, bad size or option) */
    NS_ERR_HIP = 2,     /* a HIP runtime call failed                           */
    NS_ERR_NOMEM = 3,   /* device or host allocation failed                    */
    NS_ERR_STATE = 4,   /* call out of order (e.g. step before upload)         */
    NS_ERR_NODEV = 5,   /* no usable gfx950 device                             */
    NS_ERR_IO = 6       /* a file could not be opened, written or read         */
};

/* Parameters of one step.  Field meaning follows the method-5 constants of
 * nsphere.c:5777-5782 and the force constants of nsphere.c:364-373, 513:
 * gravity on rank i is -(vel_conv_sq*G) * (i/npts) * halo_mass / r^2. */
typedef struct ns_step_params {
    double dt;
    double G;             /* G_CONST = 4.3e-6 kpc (km/s)^2 / Msun by default       */
    double halo_mass;     /* HALO_MASS = 1e12 Msun by default                      */
    double vel_conv_sq;   /* VEL_CONV_SQ = (1.02271e-3)^2 by default               */
    double radius_tol;    /* 1e-7                                                  */
    double velocity_tol;  /* 1e-7                                                  */
    double alpha;         /* r_crit = alpha ell^2 / (vel_conv_sq G M_enc), 0.05   */
    int32_t max_subdiv;   /* 4096*4096*16                                          */
    int32_t out_type;     /* 0 coarse, 1 fine, 2 4*fine-3*coarse (default)         */
    int32_t N_taumin;     /* 10                                                    */
    int32_t identity_gravity; /* 1: gravity switched off (nsphere.c:476)           */
} ns_step_params;

/* Fill *p with the reference's method-5 values in its kpc/Myr/Msun units. */
void ns_step_params_default(ns_step_params *p);
/* Same loop constants in G = M = 1 units (vel_conv_sq = 1, dt = 1e-3). */
void ns_step_params_unit(ns_step_params *p);

/* Per-particle refinement record of the last integrate (one byte each,
 * in the rank order of that step):
 *   bits 0-5 deepest level K reached (N = 2^K subdivisions), saturating at 63;
 *   bit 6    the adaptive loop hit max_subdiv without converging;
 *   bit 7    the particle took the Levi-Civita branch. */
#define NS_LVL_MASK 0x3F
#define NS_LVL_NOCONV 0x40
#define NS_LVL_LC 0x80

/* Options for ns_gpu_set_option. */
enum {
    NS_OPT_TIEBREAK = 1,   /* 0: stable on input order, exactly as quadsort
                              (default); 1: exact ties broken by particle id    */
    NS_OPT_PROFILE = 2,    /* 1: time every kernel with HIP events (those of
                              NS_OPT_PROFILE_KERNELS) and count the force
                              evaluations                                     */
    NS_OPT_PROFILE_KERNELS = 12, /* bit k set: time kernel family k (the index
                              of its name in ns_gpu_kernel_stats); default
                              all.  Each timed launch costs two events.       */
    NS_OPT_INTEGRATOR = 3, /* 0 plain: one lane per particle, levels in turn,
                                 strict IEEE order of the reference (default);
                              1 binned: leapfrog levels 0-3 per lane; deeper
                                 leapfrog particles and every Levi-Civita
                                 particle on a whole wave, one level per lane;
                                 bit-identical to 0;
                              2 binned with the fast arithmetic form (one
                                 reciprocal per force, fused multiply-adds;
                                 the Levi-Civita wave passes from level 5
                                 take the reciprocal from the previous
                                 micro-step's, Newton-refined, or its
                                 second-order series)                         */
    NS_OPT_SORT_TILE = 6,  /* radix tile: (threads << 16) | keys per thread,
                              one of 256x16, 256x24, 512x8, 512x12 (default),
                              512x16, 1024x6                                 */
    NS_OPT_SORT_ALG = 7,   /* radix pass: 0 reduce-then-scan (per pass: tile
                              digit counts, their scan, scatter; default),
                              1 onesweep (one scatter kernel per pass with
                              decoupled lookback).  Same result.             */
    NS_OPT_LANE_LEVELS = 9,  /* binned forms: leapfrog levels 0..value-1 run one
                                lane per particle before a particle moves to a
                                wave (1..6, default 4).  Same result.         */
    NS_OPT_DEEP_BLOCKS = 10, /* binned forms: workgroups (4 waves each) of the
                                wave-per-particle kernel (1..4096, default
                                1024).  Same result.                          */
    NS_OPT_GRAPH = 8,      /* 1: ns_gpu_run captures steps into a HIP graph
                              once (per parameter set and option state) and
                              replays it: two steps per launch; when the run
                              overlaps (NS_OPT_OVERLAP), the whole run in one
                              launch for an even step count up to 128
                              (ns_gpu_graph_prepare captures it ahead), else
                              one or two steps eagerly (two when that reaches
                              the buffer parity of the graph already
                              captured) and then NS_OPT_GRAPH_STEPS
                              overlapped steps per launch; both streams'
                              work is captured and each launch ends joined.
                              Needs NS_OPT_SORT_ALG 0 and profiling off (else
                              the steps are launched one by one).  Same
                              result.                                        */
    NS_OPT_GRAPH_STEPS = 14, /* overlapped steps per graph launch: even, 2..64
                              (default 16).  Each launch ends by joining the
                              second stream, so fewer steps per launch give
                              up more overlap at launch boundaries; a run
                              replays only when it has more steps than this. */
    NS_OPT_PRECISION = 5,  /* 0: float64 state (default); 1: mixed precision --
                              float32 r, v_r, ell in HBM and 32-bit radius
                              keys (4 sort passes), float64 arithmetic in the
                              integrator and the energy.  Switching converts
                              the current state (round to nearest / exact). */
    NS_OPT_BUDGET = 4,     /* force evaluations allowed per particle and step
                              (default 2^26; 0: no cap, the reference's own
                              behaviour, bounded only by max_subdiv and its
                              1e8 tau steps); a particle that needs more is
                              stopped, flagged NS_LVL_NOCONV and counted
                              (ns_gpu_cap_hits).  Not in the reference, whose
                              Levi-Civita tau steps converge only to first
                              order and can run 2^28 subdivisions each       */
    NS_OPT_OVERLAP = 11,   /* binned forms, ns_gpu_run of two or more steps:
                              1 (default): particles whose previous step was
                              slow (flagged NS_LVL_NOCONV, at leapfrog level
                              4 or Levi-Civita level 2 or deeper -- past the
                              lane kernels -- or among the 64 innermost ranks
                              of their halo; up to 1024 per step) are
                              integrated on a second stream while the main
                              stream integrates the rest and sorts the next
                              step's radii; a merge then puts them at their
                              sorted ranks.  0: every kernel on the context's
                              stream.  Same result; with halos, per halo.
                              Not taken with tiebreak by id, per-particle
                              masses or learned binning.                     */
    NS_OPT_SORT_KEYS = 13  /* float64 state, one halo, npts >= 65536, the sort
                              of a step after an integrate: 1 (default) sorts
                              24-bit quantile keys that the integrate wrote
                              (the radius's bucket in 4096 quantiles of the
                              step's start radii, then 12 bits of position in
                              the bucket) by bucket -- one scatter into
                              per-bucket regions, then one workgroup per
                              bucket orders it by (key, radius, position) --
                              when a bucket fits a workgroup (npts up to about
                              5e7), else as 2; 2: the same keys by three 8-bit
                              passes, then each run of equal keys ordered by
                              radius; 0: the 64-bit radius keys (four passes on
                              their high halves, then each run ordered by the
                              low halves).  Same permutation.                */
};

typedef struct ns_gpu_ctx ns_gpu_ctx;

const char *ns_gpu_last_error(void);
/* Library version string and the gfx target it was built for. */
const char *ns_gpu_version(void);

int ns_gpu_ctx_create(int64_t npts, int device, ns_gpu_ctx **out);
int ns_gpu_ctx_destroy(ns_gpu_ctx *ctx);
int ns_gpu_set_option(ns_gpu_ctx *ctx, int option, int64_t value);
int64_t ns_gpu_npts(const ns_gpu_ctx *ctx);

/* particles[0..3] = r, v_r, ell, id (as double), the reference's layout;
 * particles[4] (mu) is neither read nor written (ns_gpu_permute_column
 * moves it with a reordering).  An upload keeps per-particle masses by
 * position: mass k goes with the particle uploaded at position k. */
int ns_gpu_upload(ns_gpu_ctx *ctx, double *const particles[5]);
int ns_gpu_download(ns_gpu_ctx *ctx, double *particles[5]);
/* Typed SoA host arrays (any of the outputs may be NULL). */
int ns_gpu_upload_soa(ns_gpu_ctx *ctx, const double *r, const double *v,
                      const double *ell, const uint32_t *id);
int ns_gpu_download_soa(ns_gpu_ctx *ctx, double *r, double *v, double *ell,
                        uint32_t *id);
/* Device pointers of the current state (valid until the next step). */
int ns_gpu_device_state(ns_gpu_ctx *ctx, double **r, double **v, double **ell,
                        uint32_t **id);

/* sort_particles semantics: reorder all columns by radius, in place. */
int ns_gpu_sort_particles(ns_gpu_ctx *ctx);
/* The permutation of the last reordering of the state -- ns_gpu_sort_particles,
 * or the sort of the last step of ns_gpu_step / ns_gpu_run: position k of the
 * current order holds the particle that sat at position host_out[k] before it.
 * The reference moves all five columns of particles[5][npts] by this
 * permutation (nsphere.c:8613-8647).  NS_ERR_STATE before any reordering
 * since the last upload (or after ns_gpu_autotune).  host_out: npts. */
int ns_gpu_permutation(ns_gpu_ctx *ctx, uint32_t *host_out);
/* A caller-held column that the library does not keep (the reference's mu,
 * particles[4], which sort_particles_with_alg moves with the others,
 * nsphere.c:8613-8647) reordered on the device by that same permutation:
 * out[k] = in[perm[k]].  in and out may be the same array (npts doubles). */
int ns_gpu_permute_column(ns_gpu_ctx *ctx, const double *in, double *out);
/* Allocation-free radial sort.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * semantics of ns_gpu_sort_particles (stable on radius, ties in input order,
 * the same kernels) on arrays the caller owns, without a context: device
 * arrays r, v, ell (float64) and id of n particles in, their radius order out
 * in r_out, v_out, ell_out, id_out (which must not overlap any other array).
 * The workspace is a static device array of the library sized for
 * NS_SORT_NOALLOC_MAX particles (under 32 MiB, held by every GPU the library
 * runs on until the process ends); a larger n is refused (NS_ERR_ARG).  All
 * calls on one GPU share that array, so calls that overlap in time on one
 * GPU race.  Queued on stream (a hipStream_t, NULL for the null stream) of
 * the current device; returns without waiting.  Calls no allocator. */
#define NS_SORT_NOALLOC_MAX (1 << 20)
int ns_gpu_sort_particles_noalloc(const double *r, const double *v, const double *ell, const uint32_t *id,
                                  double *r_out, double *v_out, double *ell_out, uint32_t *id_out, int64_t n,
                                  void *stream);
/* inverse_map[id] = rank for the current order (host_out may be NULL to
 * keep it on the device only). */
int ns_gpu_inverse_map(ns_gpu_ctx *ctx, int32_t *host_out);
/* The per-particle loop at frozen rank (no sort): particles must already be
 * in radius-rank order, as after ns_gpu_sort_particles. */
int ns_gpu_integrate(ns_gpu_ctx *ctx, const ns_step_params *p);
/* One method-5 step: sort, then integrate at the new ranks.  Asynchronous:
 * returns once the work is queued on the context's stream. */
int ns_gpu_step(ns_gpu_ctx *ctx, const ns_step_params *p);
/* nsteps steps back to back, queued without host synchronisation (with
 * NS_OPT_OVERLAP the context's stream waits for the second stream before the
 * call returns, so the context's stream orders all later work). */
int ns_gpu_run(ns_gpu_ctx *ctx, const ns_step_params *p, int64_t nsteps);
/* With NS_OPT_GRAPH, an overlapped ns_gpu_run of an even number of steps up
 * to 128 replays one graph of the whole run (one join, at its end).  This
 * captures that graph for a run of nsteps steps from the current state,
 * without running anything, so that the run itself only replays it; for any
 * other run it does nothing.  Not in the reference (a scheduling hint for
 * callers that time their runs). */
int ns_gpu_graph_prepare(ns_gpu_ctx *ctx, const ns_step_params *p, int64_t nsteps);
int ns_gpu_sync(ns_gpu_ctx *ctx);
/* Particle-steps stopped by the work cap (NS_OPT_BUDGET) since the context
 * was created or last reset; synchronises the context's stream. */
int ns_gpu_cap_hits(ns_gpu_ctx *ctx, uint64_t *count, int reset);

/* Per-particle masses (no reference counterpart: the reference's particles
 * all weigh halo_mass / npts and M(r) is taken from the rank, nsphere.c:513).
 * mass[k] belongs to the particle held at position k of the current state
 * (the order of the last upload, or of the last download); the masses then
 * move with the particles.  From the next integrate on, M_enc at rank i is
 * the exclusive prefix sum of the masses in rank order (a decoupled-lookback
 * scan on the device, float64 for float64 state, float32 two-sum pairs for
 * float32 state; reproducible bit for bit), the force uses it in place of
 * (i/npts) M, r_crit of rank 0 uses a tenth of its own mass, and the energy
 * uses the masses.  NULL returns to equal masses. */
int ns_gpu_set_masses(ns_gpu_ctx *ctx, const double *mass);
/* An ensemble of independent halos in one context (no reference
 * counterpart: the reference evolves one halo per run).  Particles
 * [o_h, o_h + counts[h]) of the uploaded state, o_h = counts[0] + ... +
 * counts[h-1], form halo h of total mass halo_mass[h] (the step parameters'
 * halo_mass is then not used); every step sorts each halo by radius on its
 * own (one segmented radix sort over all halos) and integrates each particle
 * with its rank and count within its halo, so that halo h evolves exactly as
 * a context holding only that halo would.  nhalos = 0 or counts = NULL ends
 * the ensemble.  Needs NS_OPT_SORT_ALG 0; excludes per-particle masses. */
int ns_gpu_set_halos(ns_gpu_ctx *ctx, int nhalos, const int64_t *counts, const double *halo_mass);

/* Learned refinement binning (no reference counterpart).  A small network,
 * 5 inputs (log r, v_r, log ell, log M_enc, log dt) -> 64 -> 64 -> 8 levels,
 * BF16 weights on the matrix cores with FP32 accumulation, trained and run by
 * this library's own kernels on the levels that the adaptive integrator
 * logged.  With a network set (ns_gpu_set_mlp), every step of the binned
 * forms predicts each particle's leapfrog level, orders the particles by it
 * (one stable radix pass), and sends particles predicted past the lane
 * levels straight to a wave; the adaptive search still starts at level 0 for
 * every particle, so results are those without the network, bit for bit. */
typedef struct ns_mlp ns_mlp;
#define NS_MLP_PARAMS 6288  /* W1[64][16] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
                               (input columns >= 5 and class rows >= 8 zero) */
int ns_mlp_create(int device, uint64_t seed, ns_mlp **out);
int ns_mlp_destroy(ns_mlp *mlp);
/* Train on n (features X[n][5] as ns_gpu_features gives them, level y[n] in
 * 0..7) for `epochs` passes of minibatches of `batch`, Adam at rate lr; the
 * inputs are normalised by the set's mean and spread and the classes
 * weighted by 1/sqrt(frequency).  loss: the last epoch's weighted mean
 * cross-entropy. */
int ns_mlp_train(ns_mlp *mlp, const float *X, const uint8_t *y, int64_t n, int epochs, int64_t batch, float lr,
                 float *loss);
/* Gradient of the weighted mean cross-entropy over (X, y) at the current
 * parameters, FP32, in the NS_MLP_PARAMS layout (a check of the training
 * kernels against a CPU restatement); loss as in ns_mlp_train. */
int ns_mlp_gradient(ns_mlp *mlp, const float *X, const uint8_t *y, int64_t n, float *grad, float *loss);
int ns_mlp_predict(ns_mlp *mlp, const float *X, int64_t n, uint8_t *level);
/* params[NS_MLP_PARAMS], norm[10] (mean[5], 1/spread[5]), class_weight[8];
 * any may be NULL. */
int ns_mlp_get_params(ns_mlp *mlp, float *params, float *norm, float *class_weight);
int ns_mlp_set_params(ns_mlp *mlp, const float *params, const float *norm, const float *class_weight);
/* Use (NULL: stop using) a network for binning; needs NS_OPT_SORT_ALG 0.  The
 * network must outlive its use by the context. */
int ns_gpu_set_mlp(ns_gpu_ctx *ctx, const ns_mlp *mlp);
/* The network's raw inputs for the current state, in the current order taken
 * as rank order (as after ns_gpu_sort_particles): host_out[npts][5]. */
int ns_gpu_features(ns_gpu_ctx *ctx, const ns_step_params *p, float *host_out);

/* ---- snapshot path (nsphere.c:6452-6533, 1295-1324) ---------------------
 *
 * One record of all_particle_data<suffix>.dat: the particle's rank (its
 * position in the rank-ordered arrays, inverse_map[id]) and its r, v_r, ell
 * rounded to float32 -- the int32 + 3 x float32 written by
 * append_all_particle_data_chunk_to_file (nsphere.c:1295-1324). */
typedef struct ns_apd_record {
    int32_t rank;
    float r, v, ell;
} ns_apd_record;

/* The records of every particle of the current state in original-id order
 * (record k for id k), as the per-dtwrite block fills them
 * (nsphere.c:6456-6471); ids must be a permutation of 0..npts-1.
 * host_out: npts records. */
int ns_gpu_snapshot_records(ns_gpu_ctx *ctx, ns_apd_record *host_out);

/* r, v_r, ell and rank of k particles given by original id (the tracked
 * particles of the reference's diagnostics); rank -1 for an id not held. */
int ns_gpu_gather_ids(ns_gpu_ctx *ctx, const uint32_t *ids, int k, double *r, double *v, double *ell,
                      int32_t *rank);

/* all_particle_data writer: each ns_apd_writer_write appends the current
 * state's records as one snapshot (npts x 16 bytes) to the file, the
 * snapshot-major layout of append_all_particle_data_chunk_to_file; snapshot s
 * lies at byte s * npts * 16 from the start of the writer's output, the
 * offset of the reference's reader (nsphere.c:1380-1385).  The records
 * kernel runs on the context's stream into one of two device buffers and the
 * copy to pinned host memory on a stream of the writer's own, so steps go on
 * while a snapshot travels; the file write of a snapshot happens on the next
 * call (or at close).  As the reference appends only full blocks of
 * block_size snapshots, close keeps whole blocks and cuts a partial last
 * one.  The file is opened for appending, as the reference opens it. */
typedef struct ns_apd_writer ns_apd_writer;
int ns_apd_writer_open(ns_gpu_ctx *ctx, const char *path, int block_size, ns_apd_writer **out);
int ns_apd_writer_write(ns_apd_writer *w);
/* Writes what is pending, cuts a partial block, frees the writer; kept
 * (may be NULL) receives the number of snapshots in the file's whole blocks.
 * The context must outlive every ns_apd_writer_write; closing needs only the
 * device, so it may also follow the context's destruction. */
int ns_apd_writer_close(ns_apd_writer *w, int64_t *kept);

/* Dynamic potential, method A (nsphere.c:7698-7738), over the current radii
 * in rank order (as after ns_gpu_sort_particles or a step's sort):
 *   psi[i] = G ((i + 1) dM / r_i + dM sum_{j > i} 1 / r_j) vel_conv_sq,
 *   psi0   = G dM (sum_j 1 / r_j) vel_conv_sq   (the value at r = 0),
 * with dM = halo_mass / npts and G, vel_conv_sq, halo_mass from p.  The sum
 * is a device suffix scan in a fixed order (the same result on every run);
 * it differs from the reference's sequential sum only in rounding.
 * psi: npts doubles (may be NULL); psi0 may be NULL. */
int ns_gpu_potential(ns_gpu_ctx *ctx, const ns_step_params *p, double *psi, double *psi0);

/* Density estimate of the current radii in rank order (as after
 * ns_gpu_sort_particles), the reference's snapshot density
 * (nsphere.c:7236-7686 with fft_gaussian_convolution, 2901-3060): enclosed
 * mass (i + 1) dM (dM = halo_mass / npts) over the strictly increasing radii,
 * decimated to min(ceil(npts / 100), 30000) points, a natural cubic spline
 * of it on 131072 points uniform in log10 r, dM/dr by central differences
 * (floored at 1e-10), smoothed by the reference's one-sided Gaussian of width
 * 0.08 (npts / 1e4)^-0.4 in log10 r as a zero-padded linear convolution, and
 * a natural cubic spline of the result evaluated at every radius.  The
 * convolution is summed directly over the taps where the Gaussian exceeds
 * 1e-300 of its peak: the reference's FFT gives the same values up to its
 * round-off, about 1e-16 of the largest density.  density: npts doubles.
 * NS_ERR_ARG where the reference skips the snapshot (fewer than three
 * decimated radii, or a degenerate range). */
typedef struct ns_density_info {
    int64_t kept;       /* strictly increasing radii                     */
    int32_t decimated;  /* spline nodes of M(r)                          */
    int32_t grid;       /* log-spaced grid points (131072)               */
    int32_t taps;       /* convolution taps summed                       */
    double sigma;       /* smoothing width in log10 r                    */
    double r_min, r_max;/* grid range                                    */
} ns_density_info;
int ns_gpu_density(ns_gpu_ctx *ctx, const ns_step_params *p, double *density, ns_density_info *info);

/* The Rank_Mass_Rad_VRad files of the current state taken as a snapshot
 * (nsphere.c:7043-7231, 7698-7850, 7911-7989), written as fprintf_bin writes
 * them (int32 for the rank, float32 for every other column):
 *   unsorted_path, a row per id:  rank, (rank + 1) dM, r, v_r, PsiA, E, ell
 *   sorted_path, by float32 radius (ties in id order):
 *                                 i, (i + 1) dM, r, v_r, PsiA, E, ell, density
 * with r, v_r, ell rounded to float32 as the snapshot stores them, rank the
 * particle's position in the rank-ordered arrays, PsiA the method-A
 * potential on the reference's linear spline through (0, PsiA0) and the
 * sorted radii forced strictly increasing by steps of 1e-5, E = PsiA -
 * ell^2 / (2 r^2) - v_r^2 / 2 and density the estimate of ns_gpu_density.
 * Either path may be NULL.  Not with halos (NS_ERR_STATE). */
int ns_gpu_rank_files(ns_gpu_ctx *ctx, const ns_step_params *p, const char *unsorted_path, const char *sorted_path);

/* ---- initial conditions (nsphere.c:2386-2487, 4908-4920) -----------------
 *
 * Initial conditions drawn on the device straight into the context's state:
 * the recipes and draws of ns_ic_generate (the same per-particle Philox
 * counters; values can differ from the host's only where the device math
 * library rounds pow / log / sin / cos differently), ids 0..npts-1. */
int ns_gpu_generate_ic(ns_gpu_ctx *ctx, int kind, uint64_t seed, double rmax, double lowl_frac,
                       double lowl_scale);
/* The reference's initial-conditions file (write_initial_conditions /
 * read_initial_conditions, nsphere.c:2386-2487): int32 npts, then for each
 * particle 5 doubles (r, |v|, ell, id, mu) in the units before the
 * conversion of nsphere.c:4908-4920.  records: npts x 5 doubles. */
int ns_ic_file_write(const char *path, int32_t npts, const double *records);
/* records NULL: *npts receives the file's particle count.  Otherwise *npts is
 * the count expected, and a file holding another count is refused
 * (NS_ERR_IO), as the reference refuses it. */
int ns_ic_file_read(const char *path, int32_t *npts, double *records);
/* Uploads such records as the state, with the reference's conversion after
 * reading: v_r = (|v| mu) conv, ell = ell conv (conv = kmsec_to_kpcmyr in the
 * reference's units, 1 for none); ids from the fourth column. */
int ns_gpu_upload_ic(ns_gpu_ctx *ctx, const double *records, double conv);

/* The enclosed mass of the last scan (npts + 2 doubles): [i] below rank i,
 * [npts] the total, [npts + 1] the rank-0 particle's mass. */
int ns_gpu_enclosed_mass(ns_gpu_ctx *ctx, double *host_out);

/* Autotuner (no reference counterpart): times `steps` steps (after one
 * untimed step) of the resident state under each candidate launch
 * configuration -- the sort's pass algorithm and tile shape (threads x keys
 * per thread; the digit width is fixed at 8 bits), the lane-per-particle
 * leapfrog levels and the wave kernel's grid -- one knob at a time, keeps the
 * fastest and restores the state, so the tuning leaves no trace but the
 * chosen options.  Every candidate gives the same result. */
typedef struct ns_tune_config {
    int32_t sort_alg, sort_threads, sort_items; /* NS_OPT_SORT_ALG / _TILE    */
    int32_t lane_levels;                        /* NS_OPT_LANE_LEVELS        */
    int32_t deep_blocks;                        /* NS_OPT_DEEP_BLOCKS        */
    int32_t sort_keys;                          /* NS_OPT_SORT_KEYS: the
                                                   digits of the sort
                                                   between steps             */
    double ms_per_step;                         /* measured (HIP events)     */
} ns_tune_config;
#define NS_TUNE_MAX 32
typedef struct ns_tune_result {
    ns_tune_config best;
    int32_t ntried;
    ns_tune_config tried[NS_TUNE_MAX];
} ns_tune_result;
int ns_gpu_autotune(ns_gpu_ctx *ctx, const ns_step_params *p, int steps, ns_tune_result *out);

/* Level records of the last integrate (npts bytes, rank order). */
int ns_gpu_levels(ns_gpu_ctx *ctx, uint8_t *host_out);
/* Shell-model energy of the current state with the rank-based mass of
 * nsphere.c:513: K = sum m (v^2 + ell^2/r^2)/2,
 * W = -sum (vel_conv_sq G) (i/npts) M m / r_(i), m = M/npts, float64 sums,
 * where i is the particle's rank among the CURRENT radii (the call sorts
 * first; the state itself is not reordered).  The reference has no total-
 * energy routine; W is the potential whose gradient is its force, so E is
 * conserved by the exact dynamics and continuous when shells cross. */
int ns_gpu_energy(ns_gpu_ctx *ctx, const ns_step_params *p, double *K, double *W);

/* Kernel timing (NS_OPT_PROFILE=1): for each kernel family, the number of
 * launches and total milliseconds measured with HIP events on the context's
 * stream since the last reset.  names[i] point at static strings. */
#define NS_MAX_KERNEL_STATS 32 /* the arrays hold this many entries */
int ns_gpu_kernel_stats(ns_gpu_ctx *ctx, int *count, const char **names,
                        int64_t *launches, double *total_ms);
int ns_gpu_kernel_stats_reset(ns_gpu_ctx *ctx);
/* Force evaluations (the work units of NS_OPT_BUDGET: the reference's
 * sequential count) spent while NS_OPT_PROFILE was on, since the last
 * ns_gpu_kernel_stats_reset: out[0] by the plain kernel or the binned forms'
 * bulk kernel, out[1] by their wave-per-particle kernel on the context's
 * stream (deep), out[2] by the late kernel of overlapped runs.  Synchronises
 * the context's stream. */
int ns_gpu_force_evals(ns_gpu_ctx *ctx, uint64_t out[3]);

/* Sum over the late kernel's launches (overlapped runs, while NS_OPT_PROFILE
 * was on, since the last ns_gpu_kernel_stats_reset) of each launch's longest
 * Levi-Civita chain: the micro-steps one wave ran one after another, which
 * bound the launch's duration.  A measurement aid with no counterpart in the
 * reference.  Synchronises both of the context's streams. */
int ns_gpu_chain_steps(ns_gpu_ctx *ctx, uint64_t *out);

/* The floor of those chains (a measurement aid with no counterpart in the
 * reference): one wave alone on `device` running the fast Levi-Civita
 * micro-steps of the late kernel's long chains (one with a Newton-refined
 * reciprocal, then three with the series one) micro_steps times (1e3 .. 2^34,
 * rounded down to a multiple of 4) on a scalar trip counter; ns_per_step
 * receives its duration per micro-step.  Run it while the device is otherwise
 * idle. */
int ns_gpu_chain_floor(int device, int64_t micro_steps, double *ns_per_step);

/* Seeded initial conditions in G = M = a = 1 units (host, deterministic per
 * particle through a Philox-4x32-10 counter stream keyed by seed):
 * kind 0 Plummer (inverse-CDF radii, Aarseth-Henon-Wielen speeds),
 * kind 1 Hernquist (inverse-CDF radii, Gaussian velocities with the isotropic
 * Jeans dispersion), kind 2 half Plummer (even ids) half Hernquist (odd ids).
 * Radii beyond rmax are redrawn; a fraction lowl_frac of particles (Bernoulli
 * pick) get ell scaled by lowl_scale.  id[i] = i. */
int ns_ic_generate(int kind, int64_t n, uint64_t seed, double rmax, double lowl_frac,
                   double lowl_scale, double *r, double *v, double *ell, uint32_t *id);
/* Philox-4x32-10 block (exposed for known-answer tests). */
void ns_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
