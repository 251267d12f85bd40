// ns_kernels.h -- launch interface of the integrate / gather / energy kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nsphere_gpu.h"

namespace ns {

// Everything the fused step kernel needs, passed by value.  State arrays are
// float64 (keys 64-bit) or, in the mixed-precision form, float32 (keys
// 32-bit); the kernels are instantiated for each.
struct StepArgs {
    const void *r0, *v0, *l0;    // state in the previous rank order
    const uint32_t *id0;
    const uint32_t *perm;        // new rank -> previous position (GATHER form)
    void *r1, *v1, *l1;          // state in the new rank order
    uint32_t *id1;
    void *keys;                  // radius keys of the new r (may be null)
    // float64 state (else null): the keys are 24-bit quantile keys on this
    // table of the start radii (quantile_key; the key sentinel is all ones)
    // instead of two-plane radius keys
    const QEntry *qtab;
    uint32_t q_stride;
    uint8_t *lvl;                // level records (may be null)
    unsigned long long *budget_hits;
    int64_t n;
    double dt, G, halo_mass, vel_conv_sq, radius_tol, velocity_tol, alpha;
    int64_t max_subdiv;
    int64_t budget;              // force evaluations allowed per particle per step
    int32_t out_type, N_taumin, identity_gravity;
    // binned integrator: queues of ranks handed between its kernels
    // ensembles (nseg > 0): ranks [seg_off[h], seg_off[h + 1]) are halo h, of
    // total mass seg_mass[h]; rank and npts in the force are the halo's own
    const int64_t *seg_off;
    const double *seg_mass;
    int32_t nseg;
    // per-particle masses (else null: M_enc from the rank, as the reference):
    // menc[i] = mass enclosed below rank i, menc[n + 1] = mass of rank 0
    const double *menc;
    // learned binning (else null): bulk thread t takes rank order[t], the
    // ranks in stable order of predicted level
    const uint32_t *order;
    uint32_t *q_lf;              // leapfrog particles past cap_N (wave per particle)
    uint32_t *q_lc;              // Levi-Civita particles (lane per particle first)
    uint32_t *q_lcw;             // Levi-Civita particles whose tau steps need more
                                 // than the lane levels (wave per particle)
    uint32_t *q_count;           // [0] length of q_lf, [1] of q_lc, [2] the deep
                                 // kernel's ticket, [3] length of q_lcw
    bool q_cleared;              // q_count already cleared for this step (launch_step_prep)
    int64_t cap_N;               // deepest leapfrog N tried lane-per-particle
    int32_t deep_blocks;         // grid of the wave-per-particle kernel (256 threads each)
    // profiling only (else null): force evaluations spent, one slot per
    // wave of a launch (slot = global wave index; each wave adds its own with
    // a plain read-modify-write, launches being stream-ordered), in rows of
    // work_stride slots: row 0 plain or bulk, 1 deep, 2 late (side stream),
    // then the kChainWords chain counters
    unsigned long long *work;
    int64_t work_stride;
    // overlapped run (else null): every result that slow_region places
    // appends its rank (the late kernel: its slot | kDeferBit) and its
    // record word (slow_word) to that region of slow_list / slow_rec, counted
    // in slow_count[slow_region_ctl(region)] (the overlap's control words;
    // the counts go on past the capacities).  In that mode
    // a perm entry with kDeferBit set marks a rank that the late kernel
    // integrates on the side stream: the bulk kernel only carries its ell and
    // id and writes the key sentinel (all ones, above every radius key).
    uint32_t *slow_list;
    uint32_t *slow_rec;
    uint32_t *slow_count;
};

constexpr int kWorkKernels = 3;
// After the work rows (profiling only): the longest Levi-Civita chain of the
// running late launch (micro-steps, atomic max) and the sum of the longest
// chains of the launches before it, folded in by the plan kernel that
// precedes each late launch on the side stream.
constexpr int kChainMax = 0;
constexpr int kChainSum = 1;
constexpr int kChainWords = 2;
constexpr int kQueueCounters = 4;
constexpr uint32_t kDeferBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;
constexpr int kSlowCap = 1024;   // particles deferred per step
// Which particles are likely to need a wave of their own in the next step.
// Class 2: the chain hit the cap, or it went past the lane kernels -- a
// leapfrog level of kSlowLevel or more (past the bulk kernel's levels 0-3)
// or a Levi-Civita tau step at level kSlowLevelLc or more (past the lane
// kernel's levels 0-1) -- so that the previous step's wave particles leave
// the context's stream for the late kernel; a particle plunging towards the
// centre climbs through these levels before it turns into a capped
// Levi-Civita chain.  Class 1: otherwise, the particle sits among the
// innermost kCoreRanks ranks of its halo, where the plunges end.  Class 3:
// otherwise, the particle-step spent kSlowSpent force evaluations or more at
// low levels (a Levi-Civita particle whose tau steps shrink near pericentre
// takes many on one lane; such particles often need a wave in the next
// step).  The deferred particles are taken class 2 first, then class 1, then
// class 3, costliest first within a class.
constexpr int kSlowLevel = 4;
constexpr int kSlowLevelLc = 2;
constexpr int kSlowLevelHigh = 8;     // leapfrog levels from here on: the long class-2 chains
constexpr uint64_t kSlowSpent = 256;  // a Levi-Civita lane particle-step averages about 200
constexpr int64_t kCoreRanks = 64;
// The slow list is cut into regions, each with its own capacity and count
// (a control word of the overlap), so that no kind of candidate crowds out
// another: the first come keep their places, and a workload can produce tens
// of thousands of one kind per step while the chains that matter most -- the
// capped ones and the deep ones, which end last -- append last.
//   region 0, class 2: capped (not converged)
//   region 1, class 2: Levi-Civita, leapfrog level kSlowLevelHigh or more,
//             or a falling leapfrog particle about to turn Levi-Civita (soon)
//   region 2, class 2: the other leapfrog levels
//   region 3, class 1; region 4, class 3.
constexpr int kSlowRegions = 5;
constexpr int kSlowLists = 7 * kSlowCap;  // entries of all regions
__host__ __device__ constexpr uint32_t slow_region_first(int g)
{
    return (uint32_t)kSlowCap * (g == 0 ? 0u : g == 1 ? 1u : g == 2 ? 3u : g == 3 ? 5u : 6u);
}
__host__ __device__ constexpr uint32_t slow_region_cap(int g) { return (uint32_t)kSlowCap * (g == 1 || g == 2 ? 2u : 1u); }
__host__ __device__ constexpr int slow_region_ctl(int g) { return g == 0 ? 6 : g == 1 ? 0 : g == 2 ? 7 : g == 3 ? 3 : 5; }
__host__ __device__ constexpr int slow_region_class(int g) { return g <= 2 ? 2 : (g == 3 ? 1 : 3); }
static_assert(slow_region_first(4) + slow_region_cap(4) == (uint32_t)kSlowLists, "regions fill the list");
// The region of a result (-1: not slow).
__host__ __device__ inline int slow_region(int rec, int64_t local_rank, uint64_t spent = 0, bool soon = false)
{
    const int level = rec & NS_LVL_MASK;
    const bool lc = (rec & NS_LVL_LC) != 0;
    if (rec & NS_LVL_NOCONV)
        return 0;
    if (level >= (lc ? kSlowLevelLc : kSlowLevel))
        return (lc || level >= kSlowLevelHigh || soon) ? 1 : 2;
    if (local_rank < kCoreRanks)
        return 3;
    return spent >= kSlowSpent ? 4 : -1;
}
// A slow-list entry's record word: the level record in bits 0-7 and, for a
// particle-step that spent kSlowSpent force evaluations or more, 1 +
// log2(spent / kSlowSpent) (at most 48) in bits 8-13, which orders class 3
// when the deferred particles are chosen.
__host__ __device__ inline uint32_t slow_word(int rec, uint64_t spent)
{
    uint32_t bin = 0;
    for (uint64_t s = kSlowSpent; spent >= s && bin < 48u; s <<= 1)  // s stays below 2^57
        bin++;
    return (uint32_t)(rec & 0xFF) | (bin << 8);
}
// Append (tag, word) to region g of the slow list (entries list / words,
// count ctl[slow_region_ctl(g)], which goes on past the capacity; entries
// past it are dropped) for every lane with pred set.  One atomic per wave,
// as a lane per particle may append tens of thousands a launch (on one
// address they would serialise).  Must be reached by all active lanes.
__device__ inline void slow_push(uint32_t *ctl, uint32_t *list, uint32_t *words, int g, bool pred, uint32_t tag,
                                 uint32_t word)
{
    const uint64_t m = __ballot(pred);
    if (m == 0)
        return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(ctl + slow_region_ctl(g), (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    const uint64_t below = lane ? (m << (64 - lane)) : 0ull;  // lanes of m under this one
    const uint32_t k = base + (uint32_t)__popcll(below);
    if (pred && k < slow_region_cap(g)) {
        list[slow_region_first(g) + k] = tag;
        words[slow_region_first(g) + k] = word;
    }
}
// slots per row: enough for the waves of any integrate launch over n particles
int64_t work_stride_for(int64_t n);

// Integrator forms (NS_OPT_INTEGRATOR).
enum IntegratorForm : int {
    FORM_PLAIN = 0,        // one lane per particle, sequential levels, strict IEEE
    FORM_BINNED = 1,       // bulk lanes + deep queue with wave-parallel levels, strict IEEE
    FORM_BINNED_FAST = 2   // as BINNED with the fast arithmetic form
};

// Workspace of the enclosed-mass scan (ns_scan.hip).
struct MassScan {
    uint64_t *words;   // mass_scan_words(n) lookback words (zeroed at allocation)
    uint32_t *ctl;     // [0] tile ticket, [1] lookback epoch (device-owned)
    double scale;      // fixed-point scale, a power of two with total mass * scale < 2^62
};
int64_t mass_scan_words(int64_t n);
// menc[i] (i < n) = sum of the masses of ranks < i, menc[n] = total,
// menc[n + 1] = mass of rank 0; masses are mass_in[perm[i]] (perm may be
// null), copied to mass_out (may be null) in rank order.
template <typename M>
hipError_t launch_mass_scan(MassScan &w, const M *mass_in, const uint32_t *perm, M *mass_out, double *menc,
                            int64_t n, hipStream_t stream);

} // namespace ns
struct ns_mlp;
namespace ns {
// Learned binning (ns_mlp.hip): predicted level (0..7) of every rank of the
// step described by a (gathering through a.perm when gather), as 32-bit
// labels; raw features of the current state in rank order (n x 5 floats).
hipError_t launch_mlp_predict(const StepArgs &a, bool gather, bool f32, const float *P, const float *norm,
                              uint32_t *pred, hipStream_t stream);
hipError_t launch_mlp_features(const StepArgs &a, bool f32, float *out, hipStream_t stream);
const float *mlp_params(const ns_mlp *m);
const float *mlp_norm(const ns_mlp *m);
int mlp_device(const ns_mlp *m);

struct Profiler;
// Times the plain form as K_INTEGRATE, the binned forms' three kernels as
// K_INT_BULK and K_INT_DEEP (prof may be null).
hipError_t launch_integrate(const StepArgs &a, bool gather, int form, bool f32, hipStream_t stream,
                            Profiler *prof);

// Overlapped run (ns_overlap.hip).  Device control words of ctl:
//   [0], [3], [5], [6], [7] the slow list's region counts of this step (all
//       kernels; slow_region_ctl)
//   [1] deferred ranks of the coming step (entries of def_pos)
//   [2] ranks deferred in the step just sorted (sorted last, by key sentinel)
//   [4] slow-list entries selected for deferral (entries of sel)
struct OverlapBufs {
    uint32_t *perm;      // n: the merged order, kDeferBit on deferred ranks
    uint32_t *slow;      // kSlowLists: StepArgs::slow_list
    uint32_t *slow_rec;  // kSlowLists: the record word of each entry (slow_word)
    uint32_t *sel;       // kSlowCap: the slow-list entries deferred, costliest first
    uint32_t *def_pos;   // kSlowCap: deferred ranks, costliest first
    uint32_t *def_src;   // kSlowCap: their positions in the previous order
    uint32_t *late_h, *late_t, *late_slot;  // kSlowCap each: the late particles in (halo, key, rank) order
    uint64_t *ins;       // kSlowCap: (halo << 32 | insertion point) of the late particles, ascending
    uint32_t *found;     // 2 kSlowCap: search results of the plan
    uint32_t *ctl;       // 8 words
};
// Halos (seg_off, nseg > 0): ranks [seg_off[h], seg_off[h + 1]) are halo h,
// sorted as a segment of their own; nseg == 0: one array.
// Clears the control words; with lvl (the level records of the current
// order) the slow list is seeded from them, else left empty.
hipError_t overlap_reset(const OverlapBufs &o, const uint8_t *lvl, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream);
// The merge of one overlapped step, in three launches.  overlap_order, once
// the previous step's integrate kernels (both streams) have ended -- it does
// not need this step's sort: the late particles (the previous step's
// deferred ranks, old ranks def_pos) in (halo, key, old rank) order, and the
// slow-list entries to defer next.  overlap_plan, after the sort of r (state
// type S, keys of deferred ranks = sentinel) into sorted: the deferred ranks
// of the coming step (the previous step's slow list at their new ranks) in
// def_pos / def_src / ctl[1], and the late ranks' merged positions in perm --
// all the late kernel reads.  overlap_place: the early elements at their
// merged positions in perm; with the deferred ranks then flagged
// (launch_step_prep), what the bulk kernel reads.  chain (profiling, else null): the kChainWords counters,
// whose running maximum overlap_plan folds into the sum.
hipError_t overlap_order(const OverlapBufs &o, const void *r, bool f32, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream);
hipError_t overlap_plan(const OverlapBufs &o, const uint32_t *sorted, const void *r, bool f32, const int64_t *seg_off,
                        int nseg, int64_t n, unsigned long long *chain, hipStream_t stream);
hipError_t overlap_place(const OverlapBufs &o, const uint32_t *sorted, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream);
// The late kernel: rank def_pos[j], gathered from position def_src[j] of the
// previous order, integrated on one wave per j < *count (a.keys unused), on
// the side stream.
hipError_t launch_late(const StepArgs &a, int form, bool f32, const uint32_t *def_pos, const uint32_t *def_src,
                       const uint32_t *count, hipStream_t stream);
// One wave running the late kernel's fast Levi-Civita micro-step `steps` times.
hipError_t launch_chain_floor(int64_t steps, double *out, hipStream_t stream);
// What a step needs between its placement and its integrate, in one launch
// of independent parts: the quantile table of the start radii r0 in the
// order of perm (rank -> position, kDeferBit ignored; null: r0 is in rank
// order; n >= kQMinN) when q is given; the deferred ranks def_pos[j < *def_count]
// flagged (kDeferBit) in flag_perm when given; the binned integrator's queue
// counters q_count cleared when given.
hipError_t launch_step_prep(const double *r0, const uint32_t *perm, int64_t n, QEntry *q, uint32_t *q_count,
                            const uint32_t *def_pos, const uint32_t *def_count, uint32_t *flag_perm,
                            hipStream_t stream);
// Radius keys of the deferred ranks of the last step (replacing the sentinel).
hipError_t overlap_fix_keys(const OverlapBufs &o, const void *r, bool f32, void *keys, int64_t n,
                            const QEntry *qtab, uint32_t q_stride, hipStream_t stream);
template <typename S>
hipError_t launch_gather_state(const S *r0, const S *v0, const S *l0, const uint32_t *id0, const uint32_t *perm,
                               S *r1, S *v1, S *l1, uint32_t *id1, int64_t n, hipStream_t stream);
hipError_t launch_inverse_map(const uint32_t *id, int32_t *inv, int64_t n, hipStream_t stream);
// out[i] = perm[i] without kDeferBit; out[i] = in[perm[i]] (kDeferBit ignored).
hipError_t launch_perm_copy(const uint32_t *perm, uint32_t *out, int64_t n, hipStream_t stream);
hipError_t launch_gather_column(const double *in, const uint32_t *perm, double *out, int64_t n, hipStream_t stream);
// Snapshot path (ns_io.hip): all_particle_data records in id order (16 bytes
// each) and a gather of (r, v, ell, position) by id through inv.
hipError_t launch_apd_records(const void *r, const void *v, const void *l, bool f32, const uint32_t *id, int64_t n,
                              void *rec, hipStream_t stream);
hipError_t launch_gather_ids(const void *r, const void *v, const void *l, bool f32, const int32_t *inv, int64_t n,
                             const uint32_t *ids, int k, double *out, hipStream_t stream);
// Dynamic potential (ns_scan.hip): psi per rank from the suffix sum of 1/r;
// work holds potential_work_doubles(n) doubles, the last the sum of all 1/r.
hipError_t launch_potential(const void *r, bool f32, int64_t n, double dm, double g, double vcs, double *psi,
                            double *work, hipStream_t stream);
int64_t potential_work_doubles(int64_t n);
// Density estimate (ns_density.hip) of radii in rank order: out[n] doubles.
hipError_t density_estimate(const void *r_dev, bool f32, int64_t n, double halo_mass, double *out,
                            ns_density_info *info, hipStream_t stream);
int64_t density_grid_size();
struct SortWork;
// Rank_Mass_Rad_VRad files of the current state (ns_rank.hip); io_failed is
// set when a file cannot be written.
hipError_t rank_files(SortWork &sw, const void *r, const void *v, const void *l, bool f32, const uint32_t *id,
                      int64_t n, double G, double vcs, double halo_mass, const char *unsorted_path,
                      const char *sorted_path, hipStream_t stream, int *io_failed);
// Initial conditions (ns_io.hip): drawn on the device, or converted from an
// initial-conditions file's records (5 doubles per particle).
hipError_t launch_generate_ic(int kind, uint64_t seed, double rmax, double lowl_frac, double lowl_scale, int64_t n,
                              double *r, double *v, double *l, uint32_t *id, hipStream_t stream);
hipError_t launch_ic_records(const double *rec, int64_t n, double conv, double *r, double *v, double *l,
                             uint32_t *id, hipStream_t stream);
int energy_blocks(int64_t n);
// mass (rank order via perm, may be null: equal masses m) and menc (may be
// null: M_enc from the rank) select the per-particle-mass energy; seg_off /
// seg_mass (nseg > 0) an ensemble of halos, each with its own ranks and mass.
template <typename S>
hipError_t launch_energy(const S *r, const S *v, const S *ell, const uint32_t *perm, const S *mass,
                         const double *menc, const int64_t *seg_off, const double *seg_mass, int nseg, int64_t n,
                         double m, double gpre, double halo_mass, int idg, double *partial, double *out,
                         hipStream_t stream);
// Element-wise precision conversion (float64 <-> float32, round to nearest).
template <typename D, typename S>
hipError_t launch_convert(const S *src, D *dst, int64_t n, hipStream_t stream);

} // namespace ns
