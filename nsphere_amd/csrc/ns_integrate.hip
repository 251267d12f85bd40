// ns_integrate.hip -- the per-particle loop of the method-5 step
// (nsphere.c:5796-5838) fused with the permutation gather, plus the gather
// of sort_particles, the inverse map (nsphere.c:5789-5794) and the float64
// energy reduction.
//
// Every form gathers particle perm[i] into rank i (the previous step's order
// is nearly sorted, so the gather is nearly sequential), writes ell and id in
// the new rank order, integrates (r, v_r) at frozen rank i and writes r, v,
// the radius key of the new r (the next sort's input) and a level record.
// State is float64, or float32 in the mixed-precision form: loaded values
// are widened to float64, all arithmetic is float64, results are rounded to
// the state type once, and the key is taken from the stored value.
//
// FORM_PLAIN: one lane per particle runs the whole adaptive loop.  Its cost
//   is set by the deepest particle's sequential chain.
// FORM_BINNED(_FAST): a bulk kernel runs levels 0..3 lane-per-particle and
//   queues the rest; a deep kernel then runs each queued leapfrog particle on
//   a whole wave with one adaptive level per lane (the levels of one particle
//   are independent passes from the same start); the Levi-Civita particles
//   run lane-per-particle while their tau steps stay at the lowest levels,
//   and on a whole wave, one level per lane, when one needs more.
//   Results are identical to the plain form (strict arithmetic) because every
//   level is the same computation.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"
#include "ns_physics.h"

namespace ns {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxDeepBlocks = 4096;  // bound of StepArgs::deep_blocks (work-slot sizing)

__device__ inline uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Append `value` for every lane with pred set; one atomic per wave.  Must be
// reached by all active lanes of the wave.
__device__ inline void wave_push(uint32_t *q, uint32_t *count, bool pred, uint32_t value)
{
    const uint64_t m = __ballot(pred);
    if (m == 0)
        return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (pred)
        q[base + (uint32_t)__popcll(m & lanemask_lt())] = value;
}

// Profiling: adds the wave's force evaluations (every lane's `units`) to the
// wave's own slot of the kernel's row.  Every lane of the wave must call it,
// once, at the end of the kernel.
__device__ inline void add_work(const StepArgs &a, int kernel, unsigned long long units)
{
    if (!a.work)
        return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        units += __shfl_xor(units, off, 64);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0 && units != 0)
        a.work[kernel * a.work_stride + wave] += units;
}

// Next queue position for the calling wave: one atomic per wave on a ticket,
// so that waves which drew cheap particles take the next ones (a static
// stride could stack two long chains on one wave).  Wave-uniform result.
__device__ inline uint32_t wave_ticket(uint32_t *ticket)
{
    uint32_t q = 0;
    if ((threadIdx.x & 63) == 0)
        q = atomicAdd(ticket, 1u);
    return __shfl(q, 0, 64);
}

// The sort key of new radius rs at rank i: a quantile key when the step
// has a table (searched in the kernel's staged buckets w first), else the
// two-plane radius key.
__device__ inline void write_key(const StepArgs &a, int64_t i, double rs, const QWin &w)
{
    if (a.qtab)
        static_cast<uint32_t *>(a.keys)[i] = quantile_key(a.qtab, a.q_stride, i, radius_key(rs), w);
    else
        store_key(a.keys, a.n, i, radius_key(rs));
}
__device__ inline void write_key(const StepArgs &a, int64_t i, float rs, const QWin &)
{
    store_key(a.keys, a.n, i, radius_key(rs));
}

// Quantile buckets staged around the first rank of a block whose threads
// take consecutive ranks (a block's particles move a few buckets a step):
// kQWin buckets from kQWin / 2 - 1 below the first rank's bucket.  None
// without a table or when threads take ranks in another order.  Every thread
// of the block calls it; the block synchronises before the first search.
constexpr int kQWin = 64;
__device__ inline QWin stage_rank_window(const StepArgs &a, int64_t first, uint64_t *sk, uint64_t *sms)
{
    if (!a.qtab || a.order)
        return QWin();
    int w0 = (int)(first / (int64_t)a.q_stride) - kQWin / 2 + 1;
    w0 = w0 < 0 ? 0 : (w0 > kQBuckets - kQWin ? kQBuckets - kQWin : w0);
    return qwin_stage(a.qtab, w0, kQWin, sk, sms);
}

// slow_tag: what the slow list records for a slow result (the rank, or the
// late kernel's slot | kDeferBit); local: the rank within its halo; spent:
// the particle-step's force evaluations, where the kernel counts them; soon:
// a leapfrog particle about to turn Levi-Civita (lc_soon).
template <typename S>
__device__ inline void write_result(const StepArgs &a, int64_t i, double r, double v, int rec, bool hit,
                                    uint32_t slow_tag, int64_t local, const QWin &qw = QWin(),
                                    uint64_t spent = 0, bool soon = false)
{
    const S rs = (S)r;
    static_cast<S *>(a.r1)[i] = rs;
    static_cast<S *>(a.v1)[i] = (S)v;
    if (a.keys)
        write_key(a, i, rs, qw);
    if (a.lvl)
        a.lvl[i] = (uint8_t)rec;
    if (hit)
        atomicAdd(a.budget_hits, 1ull);
    if (a.slow_list) {  // the same for every lane
        const int g = slow_region(rec, local, spent, soon);
        const uint32_t word = slow_word(rec, spent);
#pragma unroll
        for (int r = 0; r < kSlowRegions; r++)
            slow_push(a.slow_count, a.slow_list, a.slow_rec, r, g == r, slow_tag, word);
    }
}

// Previous position of rank i (perm entries may carry kDeferBit).
template <bool GATHER>
__device__ inline int64_t source_of(const StepArgs &a, int64_t i)
{
    return GATHER ? (int64_t)(a.perm[i] & kIndexMask) : i;
}

template <bool GATHER, typename S>
__device__ inline void load_particle(const StepArgs &a, int64_t i, double &r, double &v, double &ell)
{
    const int64_t src = source_of<GATHER>(a, i);
    r = (double)static_cast<const S *>(a.r0)[src];
    v = (double)static_cast<const S *>(a.v0)[src];
    ell = (double)static_cast<const S *>(a.l0)[src];
}

template <bool GATHER, typename S>
__device__ inline void carry_ell_id(const StepArgs &a, int64_t i, double ell)
{
    if (GATHER) {
        static_cast<S *>(a.l1)[i] = (S)ell;
        a.id1[i] = a.id0[source_of<GATHER>(a, i)];
    }
}

template <bool GATHER, bool FAST, typename S>
__global__ __launch_bounds__(kThreads) void k_integrate_plain(StepArgs a)
{
    const int64_t first = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * blockDim.x;
    const int64_t i = first + threadIdx.x;
    __shared__ uint64_t s_qk[kQWin + 1], s_qms[kQWin];
    const QWin qw = stage_rank_window(a, first, s_qk, s_qms);
    __syncthreads();
    unsigned long long spent = 0;
    if (i < a.n) {
        double r, v, ell;
        load_particle<GATHER, S>(a, i, r, v, ell);
        carry_ell_id<GATHER, S>(a, i, ell);
        const RankMass rm = rank_mass(a, i);
        const Phys ph = make_phys(a, i, rm, ell);
        Budget bud{a.budget, false};
        double r_new, v_new;
        int rec;
        if (takes_levi_civita(a, rm, r, ell))
            rec = adaptive_lc<FAST>(ph, a, r, v, r_new, v_new, bud);
        else
            rec = adaptive_leap<FAST>(ph, a, r, v, r_new, v_new, bud, 0);
        write_result<S>(a, i, r_new, v_new, rec, bud.hit, (uint32_t)i, rm.k.i, qw);
        spent = (unsigned long long)(a.budget - bud.left);
    }
    add_work(a, 0, spent);
}

// Particles per thread of the bulk kernel: a thread issues the permutation
// loads of all of them, then all their state gathers, before it integrates
// any, so that a wave has that many dependent gathers in flight at once.
constexpr int kBulkPPT = 2;

template <bool GATHER, bool FAST, typename S>
__global__ __launch_bounds__(kThreads) void k_integrate_bulk(StepArgs a)
{
    const int64_t first = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * (kThreads * kBulkPPT);
    const int64_t base = first + threadIdx.x;
    __shared__ uint64_t s_qk[kQWin + 1], s_qms[kQWin];
    const QWin qw = stage_rank_window(a, first, s_qk, s_qms);
    int64_t idx[kBulkPPT], src[kBulkPPT];
    bool valid[kBulkPPT], defer[kBulkPPT];
#pragma unroll
    for (int k = 0; k < kBulkPPT; k++) {
        const int64_t t = base + (int64_t)k * kThreads;
        valid[k] = t < a.n;
        // learned binning: slot t holds the rank order[t], ranks in stable
        // order of predicted level, so that a wave's lanes share their
        // expected work
        idx[k] = (a.order && valid[k]) ? (int64_t)a.order[t] : t;
    }
    uint32_t pe[kBulkPPT];
#pragma unroll
    for (int k = 0; k < kBulkPPT; k++)
        pe[k] = (GATHER && valid[k]) ? a.perm[idx[k]] : 0u;
    double r[kBulkPPT], v[kBulkPPT], ell[kBulkPPT];
    uint32_t id[kBulkPPT];
#pragma unroll
    for (int k = 0; k < kBulkPPT; k++) {
        src[k] = GATHER ? (int64_t)(pe[k] & kIndexMask) : idx[k];
        // integrated by the late kernel on the side stream
        defer[k] = GATHER && a.slow_list && (pe[k] & kDeferBit);
        if (valid[k]) {
            r[k] = (double)static_cast<const S *>(a.r0)[src[k]];
            v[k] = (double)static_cast<const S *>(a.v0)[src[k]];
            ell[k] = (double)static_cast<const S *>(a.l0)[src[k]];
            id[k] = GATHER ? a.id0[src[k]] : 0u;
        }
    }
    __syncthreads();  // the staged buckets (their loads overlap the gathers)
    unsigned long long spent = 0;
#pragma unroll
    for (int k = 0; k < kBulkPPT; k++) {
        const int64_t i = idx[k];
        bool to_lc = false, to_lf = false;
        if (valid[k] && GATHER) {
            static_cast<S *>(a.l1)[i] = (S)ell[k];
            a.id1[i] = id[k];
        }
        if (valid[k] && defer[k]) {
            if (a.keys) {
                if (a.qtab)
                    static_cast<uint32_t *>(a.keys)[i] = ~0u;
                else
                    store_key(a.keys, a.n, i, ~(typename KeyOf<S>::type)0);
            }
        } else if (valid[k]) {
            const RankMass rm = rank_mass(a, i);
            if (takes_levi_civita(a, rm, r[k], ell[k])) {
                to_lc = true;
            } else {
                const Phys ph = make_phys(a, i, rm, ell[k]);
                Budget bud{a.budget, false};
                double r_new, v_new;
                const int rec = adaptive_leap<FAST>(ph, a, r[k], v[k], r_new, v_new, bud, a.cap_N);
                if (rec == kEscalate)
                    to_lf = true;
                else
                    write_result<S>(a, i, r_new, v_new, rec, bud.hit, (uint32_t)i, rm.k.i, qw, 0,
                                    a.slow_list && (rec & NS_LVL_MASK) >= kSoonLevel &&
                                        lc_soon(a, rm, r_new, v_new, ell[k]));
                spent += (unsigned long long)(a.budget - bud.left);
            }
        }
        wave_push(a.q_lc, a.q_count + 1, to_lc, (uint32_t)i);
        wave_push(a.q_lf, a.q_count + 0, to_lf, (uint32_t)i);
    }
    add_work(a, 0, spent);
}

// The Levi-Civita particles the bulk kernel queued, one lane each: the
// reference's sequential level search for every tau step while its levels
// stay below kLcLaneCapN, which holds nearly every one of them (a tau step
// of 1/(20 N_taumin) of the orbit's time scale converges at level 0); a
// particle with a deeper tau step is queued for a whole wave, which starts
// its step again from the beginning -- the same computation, so the same
// result.  Wave w takes the chunks w, w + waves, ... of 64 queue entries:
// the chunks cost about the same, and a wave with no chunk leaves at once
// without touching a shared counter (thousands of waves contending for one
// ticket cost more than the work).
constexpr int64_t kLcLaneCapN = 2;
static_assert(kLcLaneCapN == 2, "adaptive_lc_lane_fast runs levels 0 and 1, i.e. cap_N = 2");

template <bool GATHER, bool FAST, typename S>
__global__ __launch_bounds__(kThreads) void k_integrate_lc_lanes(StepArgs a)
{
    const int lane = threadIdx.x & 63;
    const uint32_t n_lc = a.q_count[1];
    const uint32_t waves = gridDim.x * (kThreads / 64);
    const uint32_t w = (blockIdx.x * kThreads + threadIdx.x) >> 6;
    // the queued ranks are scattered: the keys of every quantile bucket, for
    // blocks that take a particle
    __shared__ uint64_t s_qk[kQBuckets + 1];
    QWin qw;
    if (a.qtab && blockIdx.x * (uint32_t)kThreads < n_lc) {
        qw = qwin_stage(a.qtab, 0, kQBuckets, s_qk, nullptr);
        __syncthreads();
    }
    unsigned long long spent = 0;
    for (uint32_t base = w * 64u; base < n_lc; base += waves * 64u) {
        const uint32_t q = base + (uint32_t)lane;
        bool escalate = false;
        uint32_t i = 0;
        if (q < n_lc) {
            i = a.q_lc[q];
            double r, v, ell;
            load_particle<GATHER, S>(a, i, r, v, ell);
            const RankMass rm = rank_mass(a, i);
            const Phys ph = make_phys(a, i, rm, ell);
            Budget bud{a.budget, false};
            double r_new, v_new;
            const int rec = (FAST && a.max_subdiv >= 2 * kLcLaneCapN)
                                ? adaptive_lc_lane_fast(ph, a, r, v, r_new, v_new, bud)
                                : adaptive_lc<FAST>(ph, a, r, v, r_new, v_new, bud, kLcLaneCapN);
            if (rec == kEscalate)
                escalate = true;
            else
                write_result<S>(a, i, r_new, v_new, rec, bud.hit, i, rm.k.i, qw, (uint64_t)(a.budget - bud.left));
            spent += (unsigned long long)(a.budget - bud.left);
        }
        wave_push(a.q_lcw, a.q_count + 3, escalate, i);
    }
    add_work(a, 1, spent);
}

// One wave per queued particle: every Levi-Civita particle the lane kernel
// passed on (its tau steps' levels one per lane), then every leapfrog
// particle past cap_N (its levels one per lane), the Levi-Civita queue
// first: it holds the particles with the longest chains, which then start at
// once.  Wave w takes entry w; entries past the grid go by ticket to the
// waves that finish first, so waves without an entry never touch the
// ticket.
template <bool GATHER, bool FAST, typename S>
__global__ __launch_bounds__(kThreads) void k_integrate_deep(StepArgs a)
{
    const int lane = threadIdx.x & 63;
    const uint32_t n_lc = a.q_count[3], n_lf = a.q_count[0];
    const uint32_t waves = gridDim.x * (kThreads / 64);
    unsigned long long spent = 0;
    for (uint32_t q = (blockIdx.x * kThreads + threadIdx.x) >> 6; q < n_lc + n_lf;
         q = waves + wave_ticket(a.q_count + 2)) {
        const bool lc = q < n_lc;  // wave-uniform
        const int64_t i = lc ? a.q_lcw[q] : a.q_lf[q - n_lc];
        double r, v, ell;
        load_particle<GATHER, S>(a, i, r, v, ell);
        const RankMass rm = rank_mass(a, i);
        const Phys ph = make_phys(a, i, rm, ell);
        Budget bud{a.budget, false};
        double r_new, v_new;
        const int rec = lc ? adaptive_lc<FAST, true>(ph, a, r, v, r_new, v_new, bud)
                           : adaptive_leap_wave<FAST>(ph, a, r, v, r_new, v_new, bud);
        if (lane == 0) {
            write_result<S>(a, i, r_new, v_new, rec, bud.hit, (uint32_t)i, rm.k.i, QWin(), 0,
                            !lc && a.slow_list && (rec & NS_LVL_MASK) >= kSoonLevel &&
                                lc_soon(a, rm, r_new, v_new, ell));
            spent += (unsigned long long)(a.budget - bud.left);  // the wave's particle, once
        }
    }
    add_work(a, 1, spent);
}

// The late kernel of an overlapped run: one wave per deferred rank (block
// j takes rank def_pos[j] from position def_src[j] of the previous order),
// the whole of its step as the bulk and deep kernels would run it -- a
// Levi-Civita particle on the wave from the start, a leapfrog one
// lane-per-particle up to cap_N and then on the wave -- so the results are
// those of the other kernels bit for bit.  Its key stays the
// sentinel the bulk kernel wrote; the merge of the next step reads r.
template <bool FAST, typename S>
__global__ __launch_bounds__(64) void k_integrate_late(StepArgs a, const uint32_t *__restrict__ def_pos,
                                                       const uint32_t *__restrict__ def_src,
                                                       const uint32_t *__restrict__ count)
{
    const uint32_t j = blockIdx.x;
    if (j >= *count)
        return;
    // the late chains bound an overlapped step when they run long, and their
    // waves share SIMDs with the next step's sort and bulk kernels: their
    // instructions issue first
    __builtin_amdgcn_s_setprio(3);
    const int64_t i = def_pos[j], src = def_src[j];
    const double r = (double)static_cast<const S *>(a.r0)[src];
    const double v = (double)static_cast<const S *>(a.v0)[src];
    const double ell = (double)static_cast<const S *>(a.l0)[src];
    const RankMass rm = rank_mass(a, i);
    const Phys ph = make_phys(a, i, rm, ell);
    Budget bud{a.budget, false};
    double r_new, v_new;
    unsigned long long spent;
    int rec;
    if (takes_levi_civita(a, rm, r, ell)) {
        rec = adaptive_lc<FAST, true>(ph, a, r, v, r_new, v_new, bud);
        spent = (unsigned long long)(a.budget - bud.left);
    } else {
        rec = adaptive_leap<FAST>(ph, a, r, v, r_new, v_new, bud, a.cap_N);
        spent = (unsigned long long)(a.budget - bud.left);
        if (rec == kEscalate) {
            bud = Budget{a.budget, false};
            rec = adaptive_leap_wave<FAST>(ph, a, r, v, r_new, v_new, bud);
            spent += (unsigned long long)(a.budget - bud.left);
        }
    }
    if ((threadIdx.x & 63) == 0)
        write_result<S>(a, i, r_new, v_new, rec, bud.hit, j | kDeferBit, rm.k.i, QWin(), (uint64_t)spent,
                        !(rec & NS_LVL_LC) && (rec & NS_LVL_MASK) >= kSoonLevel && lc_soon(a, rm, r_new, v_new, ell));
    if (a.work && (threadIdx.x & 63) == 0 && bud.chain > 0)
        atomicMax(a.work + kWorkKernels * a.work_stride + kChainMax, (unsigned long long)bud.chain);
    add_work(a, 2, (threadIdx.x & 63) == 0 ? spent : 0ull);
}

// One wave running the fast Levi-Civita micro-steps of the late kernel's
// long chains (the levels from kSeriesLevel on: LcFast::step_seeded, then
// three LcFast::step_series) `steps` times (rounded down to a multiple of 4)
// on a scalar trip counter, as the wave form does; out keeps the result live,
// the elapsed-time sum included (the chains need it, so it is part of their
// floor).
__global__ __launch_bounds__(64) void k_chain_floor(int64_t steps, double *__restrict__ out)
{
    Phys ph;
    ph.gm = -1.0e-3;
    ph.l2 = 1.0e-9;
    ph.idg = false;
    LcFast lf;
    lf.init(ph, 1.0e-3 / (double)(steps > 0 ? steps : 1), 0.1, -0.2);
    repeat_uniform(steps / 4, [&]() {
        lf.step_seeded();
        lf.step_series();
        lf.step_series();
        lf.step_series();
    });
    if ((threadIdx.x & 63) == 0)
        out[0] = lf.rho + lf.v_end() + lf.t_end(0.0);
}

// Blocks [0, tb): the quantile table (tb = kQBuckets / 256 with q, else 0);
// the blocks after: the deferred ranks' flags; block 0 also clears the queue
// counters.  The parts write disjoint words (the table reads perm masked, so
// a flag set beside it changes nothing it reads).
__global__ void k_step_prep(const double *__restrict__ r0, const uint32_t *__restrict__ perm, int64_t n,
                            uint32_t stride, QEntry *__restrict__ q, uint32_t *__restrict__ q_count,
                            const uint32_t *__restrict__ def_pos, const uint32_t *__restrict__ def_count,
                            uint32_t *__restrict__ flag_perm)
{
    const int tb = q ? kQBuckets / 256 : 0;
    if (blockIdx.x == 0 && q_count && threadIdx.x < (unsigned)kQueueCounters)
        q_count[threadIdx.x] = 0;
    if ((int)blockIdx.x < tb) {
        const int b = blockIdx.x * blockDim.x + threadIdx.x;
        const bool last = b == kQBuckets - 1;
        const int64_t i0 = (int64_t)b * stride, i1 = last ? n - 1 : i0 + stride;
        const int64_t p0 = perm ? (int64_t)(perm[i0] & kIndexMask) : i0;
        const int64_t p1 = perm ? (int64_t)(perm[i1] & kIndexMask) : i1;
        const uint64_t k0 = radius_key(r0[p0]), k1 = radius_key(r0[p1]);
        q[b].k = k0;
        q[b].ms = quantile_ms(k0, k1, last);
        return;
    }
    const uint32_t j = (blockIdx.x - tb) * blockDim.x + threadIdx.x;
    if (flag_perm && j < *def_count)
        flag_perm[def_pos[j]] |= kDeferBit;
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, int n)
{
    if ((int)threadIdx.x < n)
        p[threadIdx.x] = 0;
}

template <typename S>
__global__ void k_gather_state(const S *__restrict__ r0, const S *__restrict__ v0, const S *__restrict__ l0,
                               const uint32_t *__restrict__ id0, const uint32_t *__restrict__ perm,
                               S *__restrict__ r1, S *__restrict__ v1, S *__restrict__ l1,
                               uint32_t *__restrict__ id1, int64_t n)
{
    const int64_t i = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t p = perm[i];
        r1[i] = r0[p];
        v1[i] = v0[p];
        l1[i] = l0[p];
        id1[i] = id0[p];
    }
}

// The permutation of a reordering without the overlap's defer flags.
__global__ void k_perm_copy(const uint32_t *__restrict__ perm, uint32_t *__restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = perm[i] & kIndexMask;
}

// A caller's column reordered as the state was: out[i] = in[perm[i]].
__global__ void k_gather_column(const double *__restrict__ in, const uint32_t *__restrict__ perm,
                                double *__restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[perm[i] & kIndexMask];
}

template <typename D, typename S>
__global__ void k_convert(const S *__restrict__ src, D *__restrict__ dst, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = (D)src[i];
}

__global__ void k_inverse_map(const uint32_t *__restrict__ id, int32_t *__restrict__ inv, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t k = id[i];
        if (k < (uint64_t)n)
            inv[k] = (int32_t)i;
    }
}

constexpr int kEnergyThreads = 256;

// Per-block float64 partial sums of K and W in a fixed tree order.  Rank i
// is the position in radius order: particle perm[i] (perm may be null when
// the arrays are already sorted).
template <typename S>
__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(
    const S *__restrict__ r, const S *__restrict__ v, const S *__restrict__ ell, const uint32_t *__restrict__ perm,
    const S *__restrict__ mass, const double *__restrict__ menc, const int64_t *__restrict__ seg_off,
    const double *__restrict__ seg_mass, int nseg, int64_t n, double m, double gpre, double halo_mass, int idg,
    double *__restrict__ partial)
{
    __shared__ double sk[kEnergyThreads], sw[kEnergyThreads];
    double k = 0.0, w = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kEnergyThreads;
    for (int64_t i = (int64_t)blockIdx.x * kEnergyThreads + threadIdx.x; i < n; i += stride) {
        const int64_t p = perm ? (int64_t)perm[i] : i;
        const double rr = r[p], vv = v[p], ll = ell[p];
        int64_t il = i, nl = n;  // the rank and count of i's halo
        double Ml = halo_mass, ml = m;
        if (nseg > 0) {
            int lo = 0, hi = nseg;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (seg_off[mid] <= i)
                    lo = mid;
                else
                    hi = mid;
            }
            il = i - seg_off[lo];
            nl = seg_off[lo + 1] - seg_off[lo];
            Ml = seg_mass[lo];
            ml = Ml / (double)nl;
        }
        const double mi = mass ? (double)mass[p] : ml;
        k += mi * 0.5 * (vv * vv + (ll * ll) / (rr * rr));
        if (!idg) {
            const double me = menc ? menc[i] : ((double)il / (double)nl) * Ml;
            w -= gpre * me * mi / rr;
        }
    }
    sk[threadIdx.x] = k;
    sw[threadIdx.x] = w;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sk[threadIdx.x] += sk[threadIdx.x + s];
            sw[threadIdx.x] += sw[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sk[0];
        partial[2 * blockIdx.x + 1] = sw[0];
    }
}

__global__ void k_energy_final(const double *__restrict__ partial, int nblocks, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double k = 0.0, w = 0.0;
        for (int b = 0; b < nblocks; b++) {
            k += partial[2 * b];
            w += partial[2 * b + 1];
        }
        out[0] = k;
        out[1] = w;
    }
}

template <bool GATHER, bool FAST, typename S>
hipError_t launch_form(const StepArgs &a, int form, hipStream_t stream, Profiler *prof)
{
    const dim3 grid((unsigned)((a.n + kThreads - 1) / kThreads));
    auto begin = [&](int kid) {
        if (prof)
            prof->begin(kid);
    };
    auto end = [&](int kid) {
        if (prof)
            prof->end(kid);
    };
    if (form == FORM_PLAIN) {
        begin(K_INTEGRATE);
        hipLaunchKernelGGL((k_integrate_plain<GATHER, FAST, S>), grid, dim3(kThreads), 0, stream, a);
        end(K_INTEGRATE);
        return hipGetLastError();
    }
    hipError_t e = hipSuccess;
    if (!a.q_cleared) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, a.q_count, kQueueCounters);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    begin(K_INT_BULK);
    const dim3 bulk_grid((unsigned)((a.n + kThreads * kBulkPPT - 1) / (kThreads * kBulkPPT)));
    hipLaunchKernelGGL((k_integrate_bulk<GATHER, FAST, S>), bulk_grid, dim3(kThreads), 0, stream, a);
    end(K_INT_BULK);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    const int deep_blocks = a.deep_blocks > 0 && a.deep_blocks <= kMaxDeepBlocks ? a.deep_blocks : 1024;
    begin(K_INT_LC);
    hipLaunchKernelGGL((k_integrate_lc_lanes<GATHER, FAST, S>), dim3(deep_blocks), dim3(kThreads), 0, stream, a);
    end(K_INT_LC);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    begin(K_INT_DEEP);
    hipLaunchKernelGGL((k_integrate_deep<GATHER, FAST, S>), dim3(deep_blocks), dim3(kThreads), 0, stream, a);
    end(K_INT_DEEP);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_typed(const StepArgs &a, bool gather, int form, hipStream_t stream, Profiler *prof)
{
    const bool fast = form == FORM_BINNED_FAST;
    if (gather)
        return fast ? launch_form<true, true, S>(a, form, stream, prof)
                    : launch_form<true, false, S>(a, form, stream, prof);
    return fast ? launch_form<false, true, S>(a, form, stream, prof)
                : launch_form<false, false, S>(a, form, stream, prof);
}

} // namespace

int64_t work_stride_for(int64_t n)
{
    const int64_t bulk = (n + kThreads - 1) / kThreads * (kThreads / 64);
    const int64_t deep = (int64_t)kMaxDeepBlocks * (kThreads / 64);
    return bulk > deep ? bulk : deep;
}

hipError_t launch_integrate(const StepArgs &a, bool gather, int form, bool f32, hipStream_t stream, Profiler *prof)
{
    if (a.n <= 0)
        return hipSuccess;
    return f32 ? launch_typed<float>(a, gather, form, stream, prof)
               : launch_typed<double>(a, gather, form, stream, prof);
}

hipError_t launch_chain_floor(int64_t steps, double *out, hipStream_t stream)
{
    hipLaunchKernelGGL(k_chain_floor, dim3(1), dim3(64), 0, stream, steps, out);
    return hipGetLastError();
}

hipError_t launch_late(const StepArgs &a, int form, bool f32, const uint32_t *def_pos, const uint32_t *def_src,
                       const uint32_t *count, hipStream_t stream)
{
    if (a.n <= 0)
        return hipSuccess;
    const bool fast = form == FORM_BINNED_FAST;
    const dim3 grid((unsigned)kSlowCap), block(64);
    if (f32) {
        if (fast)
            hipLaunchKernelGGL((k_integrate_late<true, float>), grid, block, 0, stream, a, def_pos, def_src, count);
        else
            hipLaunchKernelGGL((k_integrate_late<false, float>), grid, block, 0, stream, a, def_pos, def_src, count);
    } else {
        if (fast)
            hipLaunchKernelGGL((k_integrate_late<true, double>), grid, block, 0, stream, a, def_pos, def_src, count);
        else
            hipLaunchKernelGGL((k_integrate_late<false, double>), grid, block, 0, stream, a, def_pos, def_src,
                               count);
    }
    return hipGetLastError();
}

hipError_t launch_step_prep(const double *r0, const uint32_t *perm, int64_t n, QEntry *q, uint32_t *q_count,
                            const uint32_t *def_pos, const uint32_t *def_count, uint32_t *flag_perm,
                            hipStream_t stream)
{
    if (q && n < kQMinN)
        return hipErrorInvalidValue;
    const int tb = q ? kQBuckets / 256 : 0;
    const int fb = flag_perm ? kSlowCap / 256 : 0;
    const int blocks = tb + fb > 0 ? tb + fb : 1;
    hipLaunchKernelGGL(k_step_prep, dim3(blocks), dim3(256), 0, stream, r0, perm, n, q ? quantile_stride(n) : 1u, q,
                       q_count, def_pos, def_count, flag_perm);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_gather_state(const S *r0, const S *v0, const S *l0, const uint32_t *id0, const uint32_t *perm,
                               S *r1, S *v1, S *l1, uint32_t *id1, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gather_state<S>, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       stream, r0, v0, l0, id0, perm, r1, v1, l1, id1, n);
    return hipGetLastError();
}
template hipError_t launch_gather_state<double>(const double *, const double *, const double *, const uint32_t *,
                                                const uint32_t *, double *, double *, double *, uint32_t *,
                                                int64_t, hipStream_t);
template hipError_t launch_gather_state<float>(const float *, const float *, const float *, const uint32_t *,
                                               const uint32_t *, float *, float *, float *, uint32_t *, int64_t,
                                               hipStream_t);

template <typename D, typename S>
hipError_t launch_convert(const S *src, D *dst, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL((k_convert<D, S>), dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       stream, src, dst, n);
    return hipGetLastError();
}
template hipError_t launch_convert<float, double>(const double *, float *, int64_t, hipStream_t);
template hipError_t launch_convert<double, float>(const float *, double *, int64_t, hipStream_t);

hipError_t launch_perm_copy(const uint32_t *perm, uint32_t *out, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_perm_copy, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, stream, perm,
                       out, n);
    return hipGetLastError();
}

hipError_t launch_gather_column(const double *in, const uint32_t *perm, double *out, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gather_column, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, stream,
                       in, perm, out, n);
    return hipGetLastError();
}

hipError_t launch_inverse_map(const uint32_t *id, int32_t *inv, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_inverse_map, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, stream,
                       id, inv, n);
    return hipGetLastError();
}

int energy_blocks(int64_t n)
{
    int64_t b = (n + kEnergyThreads - 1) / kEnergyThreads;
    if (b > 1024)
        b = 1024;
    return b < 1 ? 1 : (int)b;
}

template <typename S>
hipError_t launch_energy(const S *r, const S *v, const S *ell, const uint32_t *perm, const S *mass,
                         const double *menc, const int64_t *seg_off, const double *seg_mass, int nseg, int64_t n,
                         double m, double gpre, double halo_mass, int idg, double *partial, double *out,
                         hipStream_t stream)
{
    const int blocks = energy_blocks(n);
    hipLaunchKernelGGL(k_energy_partial<S>, dim3(blocks), dim3(kEnergyThreads), 0, stream, r, v, ell, perm, mass,
                       menc, seg_off, seg_mass, nseg, n, m, gpre, halo_mass, idg, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, stream, partial, blocks, out);
    return hipGetLastError();
}
template hipError_t launch_energy<double>(const double *, const double *, const double *, const uint32_t *,
                                          const double *, const double *, const int64_t *, const double *, int,
                                          int64_t, double, double, double, int, double *, double *, hipStream_t);
template hipError_t launch_energy<float>(const float *, const float *, const float *, const uint32_t *,
                                         const float *, const double *, const int64_t *, const double *, int,
                                         int64_t, double, double, double, int, double *, double *, hipStream_t);

} // namespace ns
