// ns_api.hip -- context, memory and the C-ABI entry points of nsphere_amd.
//
// A context holds one GPU's copy of the particle state as SoA arrays in HBM,
// double-buffered (the step gathers from buffer `cur` into buffer `cur^1` in
// the new rank order), plus the radix-sort workspace.  Bytes per particle:
// 2 x (3 x 8 + 4) state + 2 x (8 + 4) sort ping-pong + 4 permutation + 1 level
// + 4 inverse map = 89 B (about 1.1 GB at 1.2e7 particles).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <new>

#include <unistd.h>

#include "ns_device.h"
#include "ns_internal.h"
#include "ns_kernels.h"

struct ns_gpu_ctx {
    int device = 0;
    int64_t n = 0;
    hipStream_t stream = nullptr;
    int cur = 0;
    double *r[2] = {}, *v[2] = {}, *l[2] = {};
    // mixed-precision form: float32 state (allocated on first use); the
    // float64 buffers then serve only as upload / download staging
    float *rf[2] = {}, *vf[2] = {}, *lf[2] = {};
    bool f32 = false;
    uint32_t *id[2] = {};
    uint8_t *lvl = nullptr;
    uint32_t *perm = nullptr;
    int32_t *inv = nullptr;
    // the permutation of the last reordering of the state (c->perm, or the
    // merged order of an overlapped run, whose entries carry kDeferBit);
    // null after an upload
    const uint32_t *last_perm = nullptr;
    ns::SortWork sw = {};
    double *energy_partial = nullptr;
    double *energy_out = nullptr;  // device, 2 doubles
    unsigned long long *budget_hits = nullptr;
    // per-particle masses (optional): float64 / float32 per buffer, like the
    // state; menc = enclosed mass per rank of the last scan (n + 2 doubles)
    double *md[2] = {};
    float *mf[2] = {};
    double *menc = nullptr;
    bool has_mass = false;
    ns::MassScan ms = {};
    // learned binning: the network and the rank order by predicted level
    const ns_mlp *mlp = nullptr;
    uint32_t *order = nullptr;
    uint32_t *id_order = nullptr;  // NS_OPT_TIEBREAK=1: the id order of the sort's first stage
    // ensemble of halos (nseg > 0): halo h holds ranks [seg_off[h], seg_off[h+1])
    int nseg = 0;
    int64_t *seg_off = nullptr;
    double *seg_mass = nullptr;
    unsigned long long *work = nullptr;  // force-evaluation counters (profiling)
    int64_t work_stride = 0;
    uint32_t *q_lf = nullptr, *q_lc = nullptr, *q_lcw = nullptr, *q_count = nullptr;
    // NS_OPT_OVERLAP: merge buffers, the second stream and its fork / join
    // events (allocated on first use)
    int overlap = 1;
    ns::OverlapBufs ov = {};
    hipStream_t side = nullptr;
    // fork / join events of eager steps, and a second set that only graph
    // captures record (a captured record leaves its event unusable eagerly)
    hipEvent_t ev[4] = {}, ev_cap[4] = {};
    bool uploaded = false;
    bool keys_valid = false;   // sw.keys[0] holds keys of r[cur]
    bool keys_q24 = false;     // ... quantile keys on qtab (else two-plane radius keys)
    int sort_keys = 1;         // NS_OPT_SORT_KEYS
    ns::QEntry *qtab = nullptr;
    bool levels_valid = false;
    int tiebreak = 0;
    int integrator = 0;
    int64_t budget = (int64_t)1 << 26;
    int64_t cap_N = 8;       // NS_OPT_LANE_LEVELS: levels 0..log2(cap_N) lane-per-particle
    int deep_blocks = 1024;  // NS_OPT_DEEP_BLOCKS
    ns::Profiler prof;
    // NS_OPT_GRAPH: ns_gpu_run replays captured steps (two per launch, or
    // graph_steps overlapped ones, NS_OPT_GRAPH_STEPS)
    bool use_graph = false;
    int graph_steps = 16;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    struct GraphKey {
        ns_step_params p;
        int cur, integrator, f32, nseg, has_mass, overlapped, sort_keys, graph_steps, keys_state;
        int64_t run_steps;  // a whole overlapped run in one graph (else 0)
        int64_t budget, cap_N;
        int deep_blocks;
        ns::SortConfig cfg;
    } gkey = {};
    // the host-side state a replay of the whole-run graph leaves
    struct GraphPost {
        int cur;
        bool keys_valid, keys_q24, levels_valid;
        const uint32_t *last_perm;
    } gpost = {};
};

namespace ns {

namespace {
thread_local char g_err[512] = "";
}

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }

const char *kernel_name(int id)
{
    static const char *names[K_COUNT] = {"radius_keys", "sort_zero", "sort_histogram", "sort_scatter",
                                         "gather_state", "integrate", "inverse_map", "energy",
                                         "sort_tiefix", "sort_upsweep", "sort_scan", "integrate_bulk",
                                         "integrate_deep", "mass_scan", "mlp_binning",
                                         "snapshot_records", "potential", "density",
                                         "rank_files", "integrate_late", "overlap_plan",
                                         "overlap_place", "sort_fixup", "integrate_lc",
                                         "qsort_scatter", "qsort_bucket", "overlap_order"};
    return (id >= 0 && id < K_COUNT) ? names[id] : "?";
}

hipEvent_t Profiler::take()
{
    if (npool_ > 0)
        return pool_[--npool_];
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

void Profiler::begin(int kid, hipStream_t s)
{
    if (!on || !((mask >> kid) & 1u))
        return;
    open_ = take();
    open_kid_ = kid;
    if (open_)
        (void)hipEventRecord(open_, s ? s : stream);
}

void Profiler::end(int kid, hipStream_t s)
{
    if (!on || !open_ || open_kid_ != kid)
        return;
    hipEvent_t b = take();
    if (!b)
        return;
    (void)hipEventRecord(b, s ? s : stream);
    if (npend_ == cap_) {
        int nc = cap_ ? 2 * cap_ : 256;
        Pending *np = (Pending *)realloc(pend_, (size_t)nc * sizeof(Pending));
        if (!np)
            return;
        pend_ = np;
        cap_ = nc;
    }
    pend_[npend_++] = Pending{kid, open_, b};
    open_ = nullptr;
    if (npend_ >= 4096)
        (void)fold();
}

hipError_t Profiler::fold()
{
    hipError_t err = hipSuccess;
    for (int i = 0; i < npend_; i++) {
        hipError_t e = hipEventSynchronize(pend_[i].b);
        float ms_ = 0.f;
        if (e == hipSuccess)
            e = hipEventElapsedTime(&ms_, pend_[i].a, pend_[i].b);
        if (e == hipSuccess) {
            launches[pend_[i].kid]++;
            ms[pend_[i].kid] += ms_;
        } else {
            err = e;
        }
        for (hipEvent_t ev : {pend_[i].a, pend_[i].b}) {
            if (npool_ == pool_cap_) {
                int nc = pool_cap_ ? 2 * pool_cap_ : 512;
                hipEvent_t *np = (hipEvent_t *)realloc(pool_, (size_t)nc * sizeof(hipEvent_t));
                if (!np) {
                    (void)hipEventDestroy(ev);
                    continue;
                }
                pool_ = np;
                pool_cap_ = nc;
            }
            pool_[npool_++] = ev;
        }
    }
    npend_ = 0;
    return err;
}

void Profiler::reset()
{
    (void)fold();
    for (int k = 0; k < K_COUNT; k++) {
        launches[k] = 0;
        ms[k] = 0.0;
    }
}

void Profiler::release()
{
    (void)fold();
    for (int i = 0; i < npool_; i++)
        (void)hipEventDestroy(pool_[i]);
    free(pool_);
    free(pend_);
    pool_ = nullptr;
    pend_ = nullptr;
    npool_ = pool_cap_ = npend_ = cap_ = 0;
}

} // namespace ns

using ns::set_error;

namespace {

#define HIP_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return set_error(NS_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                            \
    } while (0)

// The profiling counters: the work rows and the chain counters after them.
size_t work_bytes(const ns_gpu_ctx *c)
{
    return ((size_t)ns::kWorkKernels * (size_t)c->work_stride + ns::kChainWords) * sizeof(unsigned long long);
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        if (prev != dev)
            (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

template <typename T>
hipError_t dalloc(T **p, int64_t count)
{
    size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(T);
    return hipMalloc((void **)p, bytes);
}

void free_all(ns_gpu_ctx *c)
{
    ns::seg_table_release(c->sw.seg);
    ns::qsort_release(c->sw.qs);
    void *bufs[] = {c->r[0], c->r[1], c->v[0], c->v[1], c->l[0], c->l[1], c->id[0], c->id[1],
                    c->sw.keys[0], c->sw.keys[1], c->sw.vals[0], c->sw.vals[1], c->lvl, c->perm,
                    c->inv, c->sw.hist, c->sw.lookback, c->sw.counters, c->energy_partial,
                    c->energy_out, c->budget_hits, c->work, c->q_lf, c->q_lc, c->q_lcw, c->q_count,
                    c->rf[0], c->rf[1], c->vf[0], c->vf[1], c->lf[0], c->lf[1], c->md[0], c->md[1],
                    c->mf[0], c->mf[1], c->menc, c->ms.words, c->ms.ctl, c->seg_off, c->seg_mass,
                    c->order, c->id_order, c->ov.perm, c->ov.slow, c->ov.slow_rec, c->ov.sel, c->ov.def_pos, c->ov.def_src, c->ov.late_h,
                    c->ov.late_t, c->ov.late_slot, c->ov.ins, c->ov.found, c->ov.ctl, c->qtab};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    for (int k = 0; k < 4; k++)
        for (hipEvent_t e : {c->ev[k], c->ev_cap[k]})
            if (e)
                (void)hipEventDestroy(e);
    if (c->side)
        (void)hipStreamDestroy(c->side);
}

// Buffers, second stream and events of an overlapped run, on first use.
int ensure_overlap(ns_gpu_ctx *c)
{
    hipError_t e = hipSuccess;
    if (!c->ov.perm)
        e = dalloc(&c->ov.perm, c->n);
    uint32_t **lists[] = {&c->ov.slow, &c->ov.slow_rec};
    for (uint32_t **b : lists)
        if (e == hipSuccess && !*b)
            e = dalloc(b, ns::kSlowLists);
    uint32_t **small[] = {&c->ov.sel,    &c->ov.def_pos, &c->ov.def_src,
                          &c->ov.late_h, &c->ov.late_t,  &c->ov.late_slot};
    for (uint32_t **b : small)
        if (e == hipSuccess && !*b)
            e = dalloc(b, ns::kSlowCap);
    if (e == hipSuccess && !c->ov.ins)
        e = dalloc(&c->ov.ins, ns::kSlowCap);
    if (e == hipSuccess && !c->ov.found)
        e = dalloc(&c->ov.found, 2 * ns::kSlowCap);
    if (e == hipSuccess && !c->ov.ctl)
        e = dalloc(&c->ov.ctl, 8);
    if (e == hipSuccess && !c->side)
        e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    for (int k = 0; k < 4; k++)
        for (hipEvent_t *ev : {&c->ev[k], &c->ev_cap[k]})
            if (e == hipSuccess && !*ev)
                e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess)
        return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP, "overlap buffers (npts=%lld): %s",
                         (long long)c->n, hipGetErrorString(e));
    return NS_OK;
}

// Float32 state buffers of the mixed-precision form, allocated on first use.
int ensure_f32(ns_gpu_ctx *c)
{
    for (int b = 0; b < 2; b++) {
        float **bufs[] = {&c->rf[b], &c->vf[b], &c->lf[b]};
        for (float **p : bufs) {
            if (*p)
                continue;
            hipError_t e = dalloc(p, c->n);
            if (e != hipSuccess)
                return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP,
                                 "float32 state allocation (npts=%lld): %s", (long long)c->n, hipGetErrorString(e));
        }
    }
    return NS_OK;
}

// Float64 state of buffer b -> float32 state of buffer b (round to nearest), or back.
int state_to_f32(ns_gpu_ctx *c, int b)
{
    HIP_TRY((ns::launch_convert<float, double>(c->r[b], c->rf[b], c->n, c->stream)));
    HIP_TRY((ns::launch_convert<float, double>(c->v[b], c->vf[b], c->n, c->stream)));
    HIP_TRY((ns::launch_convert<float, double>(c->l[b], c->lf[b], c->n, c->stream)));
    return NS_OK;
}

int state_to_f64(ns_gpu_ctx *c, int b)
{
    HIP_TRY((ns::launch_convert<double, float>(c->rf[b], c->r[b], c->n, c->stream)));
    HIP_TRY((ns::launch_convert<double, float>(c->vf[b], c->v[b], c->n, c->stream)));
    HIP_TRY((ns::launch_convert<double, float>(c->lf[b], c->l[b], c->n, c->stream)));
    return NS_OK;
}

int check_ctx(const ns_gpu_ctx *c, const char *fn)
{
    if (!c)
        return set_error(NS_ERR_ARG, "%s: null context", fn);
    return NS_OK;
}

int check_params(const ns_step_params *p, const char *fn)
{
    if (!p)
        return set_error(NS_ERR_ARG, "%s: null params", fn);
    if (!(p->dt > 0.0) || p->out_type < 0 || p->out_type > 2)
        return set_error(NS_ERR_ARG, "%s: dt must be > 0 and out_type in {0,1,2}", fn);
    return NS_OK;
}

ns::StepArgs make_args(ns_gpu_ctx *c, const ns_step_params *p)
{
    ns::StepArgs a{};
    a.n = c->n;
    a.dt = p->dt;
    a.G = p->G;
    a.halo_mass = p->halo_mass;
    a.vel_conv_sq = p->vel_conv_sq;
    a.radius_tol = p->radius_tol;
    a.velocity_tol = p->velocity_tol;
    a.alpha = p->alpha;
    a.max_subdiv = p->max_subdiv;
    a.out_type = p->out_type;
    a.N_taumin = p->N_taumin;
    a.identity_gravity = p->identity_gravity;
    a.budget = c->budget;
    a.budget_hits = c->budget_hits;
    a.q_lf = c->q_lf;
    a.q_lc = c->q_lc;
    a.q_lcw = c->q_lcw;
    a.q_count = c->q_count;
    a.cap_N = c->cap_N;  // leapfrog levels lane-per-particle (default 0..3)
    a.deep_blocks = c->deep_blocks;
    a.work = c->prof.on ? c->work : nullptr;
    a.menc = c->has_mass ? c->menc : nullptr;
    a.nseg = c->nseg;
    a.seg_off = c->seg_off;
    a.seg_mass = c->seg_mass;
    a.work_stride = c->work_stride;
    return a;
}

// Per-particle mass buffers, allocated on first use.
int ensure_mass(ns_gpu_ctx *c)
{
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; b++)
        if (!c->md[b])
            e = dalloc(&c->md[b], c->n);
    if (e == hipSuccess && c->f32)
        for (int b = 0; b < 2 && e == hipSuccess; b++)
            if (!c->mf[b])
                e = dalloc(&c->mf[b], c->n);
    if (e == hipSuccess && !c->menc)
        e = dalloc(&c->menc, c->n + 2);
    if (e == hipSuccess && !c->ms.words) {
        e = dalloc(&c->ms.words, ns::mass_scan_words(c->n));
        if (e == hipSuccess)
            e = hipMemset(c->ms.words, 0, (size_t)ns::mass_scan_words(c->n) * sizeof(uint64_t));
    }
    if (e == hipSuccess && !c->ms.ctl) {
        e = dalloc(&c->ms.ctl, 2);
        if (e == hipSuccess)
            e = hipMemset(c->ms.ctl, 0, 2 * sizeof(uint32_t));
    }
    if (e != hipSuccess)
        return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP, "mass buffers (npts=%lld): %s",
                         (long long)c->n, hipGetErrorString(e));
    return NS_OK;
}

// Enclosed mass per rank for the masses of buffer `cur` taken through perm
// (null: in place); with gather the rank-ordered masses go to buffer cur^1.
int mass_scan(ns_gpu_ctx *c, const uint32_t *perm, bool gather)
{
    if (!c->has_mass)
        return NS_OK;
    const int s = c->cur, d = c->cur ^ 1;
    c->prof.begin(ns::K_MASS_SCAN);
    if (c->f32)
        HIP_TRY(ns::launch_mass_scan<float>(c->ms, c->mf[s], perm, gather ? c->mf[d] : nullptr, c->menc, c->n,
                                            c->stream));
    else
        HIP_TRY(ns::launch_mass_scan<double>(c->ms, c->md[s], perm, gather ? c->md[d] : nullptr, c->menc, c->n,
                                             c->stream));
    c->prof.end(ns::K_MASS_SCAN);
    return NS_OK;
}

void release_graph(ns_gpu_ctx *c)
{
    if (c->graph_exec)
        (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph)
        (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

ns_gpu_ctx::GraphKey graph_key(const ns_gpu_ctx *c, const ns_step_params *p, bool overlapped, int64_t run_steps = 0)
{
    ns_gpu_ctx::GraphKey k = {};
    k.run_steps = run_steps;
    k.keys_state = (c->keys_valid ? 2 : 0) | (c->keys_q24 ? 1 : 0);
    k.p = *p;
    k.cur = c->cur;
    k.overlapped = overlapped;
    k.sort_keys = c->sort_keys;
    k.graph_steps = overlapped ? c->graph_steps : 0;
    k.integrator = c->integrator;
    k.f32 = c->f32;
    k.nseg = c->nseg;
    k.has_mass = c->has_mass;
    k.budget = c->budget;
    k.cap_N = c->cap_N;
    k.deep_blocks = c->deep_blocks;
    k.cfg = c->sw.config;
    return k;
}

bool same_key(const ns_gpu_ctx::GraphKey &a, const ns_gpu_ctx::GraphKey &b)
{
    return std::memcmp(&a.p, &b.p, sizeof a.p) == 0 && a.cur == b.cur && a.integrator == b.integrator &&
           a.f32 == b.f32 && a.nseg == b.nseg && a.has_mass == b.has_mass && a.overlapped == b.overlapped &&
           a.sort_keys == b.sort_keys && a.graph_steps == b.graph_steps && a.run_steps == b.run_steps &&
           a.keys_state == b.keys_state &&
           a.budget == b.budget &&
           a.cap_N == b.cap_N && a.deep_blocks == b.deep_blocks &&
           a.cfg.alg == b.cfg.alg && a.cfg.threads == b.cfg.threads && a.cfg.items == b.cfg.items;
}

// Sort r[cur] into out (c->perm unless given; keys from the integrate
// epilogue when valid).
int sort_perm(ns_gpu_ctx *c, uint32_t *out = nullptr)
{
    if (!out)
        out = c->perm;
    if (c->tiebreak) {  // (radius, id): ids first, radius keys gathered through their order
        const int s = c->cur;
        c->prof.begin(ns::K_TIEFIX);
        const hipError_t e = ns::radix_sort_perm_ids(c->sw, c->n, c->f32 ? (const void *)c->rf[s] : (const void *)c->r[s],
                                                     c->f32, c->id[s], c->id_order, out, c->stream, nullptr);
        c->prof.end(ns::K_TIEFIX);
        HIP_TRY(e);
        c->keys_valid = false;
        return NS_OK;
    }
    if (c->keys_valid && c->keys_q24) {
        HIP_TRY(ns::radix_sort_perm_q24(c->sw, c->n, c->r[c->cur], out, c->stream, &c->prof, c->sort_keys == 1));
        c->keys_valid = false;  // the passes overwrite keys[0]
        return NS_OK;
    }
    if (!c->keys_valid) {
        c->prof.begin(ns::K_KEYS);
        if (c->f32)
            HIP_TRY(ns::launch_radius_keys(c->rf[c->cur], reinterpret_cast<uint32_t *>(c->sw.keys[0]), c->n,
                                           c->stream));
        else
            HIP_TRY(ns::launch_radius_keys(c->r[c->cur], c->sw.keys[0], c->n, c->stream));
        c->prof.end(ns::K_KEYS);
    }
    HIP_TRY(ns::radix_sort_perm(c->sw, c->n, c->f32 ? 32 : 64, out, c->stream, &c->prof));
    c->keys_valid = false;  // the passes overwrite keys[0]
    return NS_OK;
}

// Whether the integrate of a gathered step writes quantile keys.
bool quantile_applies(const ns_gpu_ctx *c)
{
    return c->sort_keys && !c->f32 && c->nseg == 0 && c->n >= ns::kQMinN;
}

// The quantile table, and the bucket sort's workspace for the sort that
// follows (allocated here, outside any capture, since they are set up
// synchronously).
int ensure_qtab(ns_gpu_ctx *c)
{
    hipError_t e = hipSuccess;
    if (!c->qtab)
        e = hipMalloc(&c->qtab, (size_t)ns::kQBuckets * sizeof(ns::QEntry));
    if (e == hipSuccess && c->sort_keys == 1)
        e = ns::qsort_prepare(c->sw.qs, c->n);
    if (e != hipSuccess)
        return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP, "quantile keys: %s",
                         hipGetErrorString(e));
    return NS_OK;
}

// Set-up of a binned integrate, one launch: the quantile table of the
// step's start radii (buffer cur, in the order of perm) when the integrate
// writes quantile keys, the deferred ranks' flags in perm (flags: an
// overlapped step, after its placement), and the queue counters cleared.
int step_prep(ns_gpu_ctx *c, ns::StepArgs &a, const uint32_t *perm, bool flags)
{
    a.qtab = nullptr;
    a.q_stride = 0;
    a.q_cleared = false;
    const bool table = quantile_applies(c) && a.keys;
    if (table)
        if (int rc = ensure_qtab(c))
            return rc;
    if (!table && !flags && c->integrator == ns::FORM_PLAIN)
        return NS_OK;  // the plain kernel uses no queues
    HIP_TRY(ns::launch_step_prep(c->r[c->cur], perm, c->n, table ? c->qtab : nullptr,
                                 c->integrator == ns::FORM_PLAIN ? nullptr : c->q_count,
                                 flags ? c->ov.def_pos : nullptr, flags ? c->ov.ctl + 1 : nullptr,
                                 flags ? c->ov.perm : nullptr, c->stream));
    if (table) {
        a.qtab = c->qtab;
        a.q_stride = ns::quantile_stride(c->n);
    }
    a.q_cleared = c->integrator != ns::FORM_PLAIN;
    return NS_OK;
}

// Arguments of the integrate kernels for the state of buffer cur: gather
// (through perm) into buffer cur^1, or in place.
ns::StepArgs state_args(ns_gpu_ctx *c, const ns_step_params *p, bool gather, const uint32_t *perm)
{
    ns::StepArgs a = make_args(c, p);
    // gather: previous order in buffer cur -> new rank order in cur^1;
    // otherwise in place (each thread reads and writes only its own index).
    const int s = c->cur, d = gather ? (c->cur ^ 1) : c->cur;
    if (c->f32) {
        a.r0 = c->rf[s];
        a.v0 = c->vf[s];
        a.l0 = c->lf[s];
        a.r1 = c->rf[d];
        a.v1 = c->vf[d];
        a.l1 = gather ? c->lf[d] : nullptr;
    } else {
        a.r0 = c->r[s];
        a.v0 = c->v[s];
        a.l0 = c->l[s];
        a.r1 = c->r[d];
        a.v1 = c->v[d];
        a.l1 = gather ? c->l[d] : nullptr;
    }
    a.id0 = c->id[s];
    a.perm = gather ? perm : nullptr;
    a.id1 = gather ? c->id[d] : nullptr;
    a.keys = c->sw.keys[0];
    a.lvl = c->lvl;
    return a;
}

int integrate(ns_gpu_ctx *c, const ns_step_params *p, bool gather)
{
    if (int rc = mass_scan(c, gather ? c->perm : nullptr, gather))
        return rc;
    ns::StepArgs a = state_args(c, p, gather, c->perm);
    const int d = gather ? (c->cur ^ 1) : c->cur;
    if (c->mlp && c->integrator != ns::FORM_PLAIN) {
        // learned binning: predicted levels (as 32-bit labels in the key
        // buffer, free until the integrator writes the next radius keys),
        // then the ranks in stable order of predicted level
        c->prof.begin(ns::K_MLP);
        HIP_TRY(ns::launch_mlp_predict(a, gather, c->f32, ns::mlp_params(c->mlp), ns::mlp_norm(c->mlp),
                                       reinterpret_cast<uint32_t *>(c->sw.keys[0]), c->stream));
        HIP_TRY(ns::radix_sort_perm_u8(c->sw, c->n, c->order, c->stream));
        c->prof.end(ns::K_MLP);
        a.order = c->order;
    }
    if (int rc = step_prep(c, a, gather ? c->perm : nullptr, false))
        return rc;
    HIP_TRY(ns::launch_integrate(a, gather, c->integrator, c->f32, c->stream, &c->prof));
    if (gather)
        c->last_perm = c->perm;
    c->cur = d;
    c->keys_valid = true;
    c->keys_q24 = a.qtab != nullptr;
    c->levels_valid = true;
    return NS_OK;
}

// Whether ns_gpu_run overlaps the late particles with the next sort.
bool overlap_applies(const ns_gpu_ctx *c, int64_t nsteps)
{
    return c->overlap && nsteps >= 2 && c->integrator != ns::FORM_PLAIN && !c->tiebreak &&
           !c->has_mass && !c->mlp && c->n <= (int64_t)ns::kIndexMask;
}

// One step of an overlapped run.  Context's stream: the sort (late keys at
// the sentinel).  Second stream, after the previous step's integrate kernels
// on the context's stream and its own late kernel: the order of the late
// particles, beside the sort; after the sort: the plan, then this step's late
// kernel -- one chain hands over to the next without a round trip through
// the context's stream.  Context's stream, after the plan: the placement of
// the early elements and the bulk and deep kernels.  ev: the fork / join
// events (sorted, merged, late, integrated), c->ev eagerly or c->ev_cap under
// a capture.
int overlapped_step(ns_gpu_ctx *c, const ns_step_params *p, const hipEvent_t *ev)
{
    const void *r = c->f32 ? (const void *)c->rf[c->cur] : (const void *)c->r[c->cur];
    HIP_TRY(hipEventRecord(ev[3], c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, ev[3], 0));
    c->prof.begin(ns::K_ORDER, c->side);
    HIP_TRY(ns::overlap_order(c->ov, r, c->f32, c->seg_off, c->nseg, c->n, c->side));
    c->prof.end(ns::K_ORDER, c->side);
    if (int rc = sort_perm(c))
        return rc;
    HIP_TRY(hipEventRecord(ev[0], c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, ev[0], 0));
    c->prof.begin(ns::K_MERGE, c->side);
    HIP_TRY(ns::overlap_plan(c->ov, c->perm, r, c->f32, c->seg_off, c->nseg, c->n,
                             c->prof.on ? c->work + ns::kWorkKernels * c->work_stride : nullptr, c->side));
    c->prof.end(ns::K_MERGE, c->side);
    HIP_TRY(hipEventRecord(ev[1], c->side));
    ns::StepArgs a = state_args(c, p, true, c->ov.perm);
    a.slow_list = c->ov.slow;
    a.slow_rec = c->ov.slow_rec;
    a.slow_count = c->ov.ctl;  // the region counts (slow_region_ctl)
    ns::StepArgs late = a;
    late.keys = nullptr;  // the sentinel stays until the next merge
    late.qtab = nullptr;
    c->prof.begin(ns::K_INT_LATE, c->side);
    HIP_TRY(ns::launch_late(late, c->integrator, c->f32, c->ov.def_pos, c->ov.def_src, c->ov.ctl + 1, c->side));
    c->prof.end(ns::K_INT_LATE, c->side);
    HIP_TRY(hipEventRecord(ev[2], c->side));
    HIP_TRY(hipStreamWaitEvent(c->stream, ev[1], 0));
    c->prof.begin(ns::K_PLACE);
    HIP_TRY(ns::overlap_place(c->ov, c->perm, c->seg_off, c->nseg, c->n, c->stream));
    c->prof.end(ns::K_PLACE);
    if (int rc = step_prep(c, a, c->ov.perm, true))
        return rc;
    HIP_TRY(ns::launch_integrate(a, true, c->integrator, c->f32, c->stream, &c->prof));
    c->last_perm = c->ov.perm;
    c->cur ^= 1;
    c->keys_valid = true;
    c->keys_q24 = a.qtab != nullptr;
    c->levels_valid = true;
    return NS_OK;
}

// Overlapped runs of an even number of steps up to this many are captured
// whole: one launch, one join at the end, no step eagerly.
constexpr int64_t kGraphRunMax = 128;

bool whole_run_graph(int64_t nsteps) { return nsteps >= 2 && nsteps <= kGraphRunMax && (nsteps & 1) == 0; }

// Capture nsteps overlapped steps (an even count, so that the double
// buffers return to the captured parity) and the join of the second stream
// into c->graph_exec (the fork / join events become the graph's
// edges between its two branches).  Capture records without running, so the
// host-side state the steps advance is put back after recording; on success
// the recorded steps leave it where it started (even parity flips, keys and
// levels valid).
int capture_overlapped(ns_gpu_ctx *c, const ns_step_params *p, const ns_gpu_ctx::GraphKey &key, int64_t nsteps)
{
    release_graph(c);
    // what a step sets up on its first call, which a capture must not meet
    HIP_TRY(ns::sort_prepare(c->sw, c->n));
    if (quantile_applies(c))
        if (int rc = ensure_qtab(c))
            return rc;
    const int cur = c->cur;
    const bool keys_valid = c->keys_valid, keys_q24 = c->keys_q24, levels_valid = c->levels_valid;
    const uint32_t *last_perm = c->last_perm;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = NS_OK;
    for (int64_t k = 0; k < nsteps && rc == NS_OK; k++)
        rc = overlapped_step(c, p, c->ev_cap);
    const hipError_t ej = hipStreamWaitEvent(c->stream, c->ev_cap[2], 0);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    c->gpost = {c->cur, c->keys_valid, c->keys_q24, c->levels_valid, c->last_perm};
    c->cur = cur;
    c->keys_valid = keys_valid;
    c->keys_q24 = keys_q24;
    c->levels_valid = levels_valid;
    c->last_perm = last_perm;
    if (rc || ej != hipSuccess || e != hipSuccess) {
        if (gr)
            (void)hipGraphDestroy(gr);
        if (rc)
            return rc;
        return set_error(NS_ERR_HIP, "ns_gpu_run: capture of overlapped steps failed: %s",
                         hipGetErrorString(ej != hipSuccess ? ej : e));
    }
    c->graph = gr;
    const hipError_t ei = hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        release_graph(c);
        return set_error(NS_ERR_HIP, "ns_gpu_run: graph instantiation failed: %s", hipGetErrorString(ei));
    }
    c->gkey = key;
    return NS_OK;
}

// nsteps >= 2 overlapped steps; the context's stream then waits for the
// second stream and the last step's late ranks get their true keys.  With
// graphs: the first step eagerly (the first two, when the captured graph
// expects the other buffer parity and the run is long enough), a join, then
// launches of c->graph_steps captured steps each (which end joined), then
// the rest eagerly.
int run_overlapped(ns_gpu_ctx *c, const ns_step_params *p, int64_t nsteps, bool graph)
{
    if (int rc = ensure_overlap(c))
        return rc;
    // the slow list of the last step, from its level records (whatever form
    // ran it): the first step defers the chains that step found slow
    HIP_TRY(ns::overlap_reset(c->ov, c->levels_valid ? c->lvl : nullptr, c->seg_off, c->nseg, c->n, c->stream));
    if (graph && whole_run_graph(nsteps)) {  // the whole run, one launch
        const ns_gpu_ctx::GraphKey key = graph_key(c, p, true, nsteps);
        if (!c->graph_exec || !same_key(key, c->gkey))
            if (int rc = capture_overlapped(c, p, key, nsteps))
                return rc;
        HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
        c->cur = c->gpost.cur;
        c->keys_valid = c->gpost.keys_valid;
        c->keys_q24 = c->gpost.keys_q24;
        c->levels_valid = c->gpost.levels_valid;
        c->last_perm = c->gpost.last_perm;
        const void *r = c->f32 ? (const void *)c->rf[c->cur] : (const void *)c->r[c->cur];
        HIP_TRY(ns::overlap_fix_keys(c->ov, r, c->f32, c->sw.keys[0], c->n, c->keys_q24 ? c->qtab : nullptr,
                                     ns::quantile_stride(c->n), c->stream));
        return NS_OK;
    }
    const int64_t G = c->graph_steps;
    int rc = overlapped_step(c, p, c->ev);
    int64_t done = 1;
    if (rc == NS_OK && graph && c->graph_exec && nsteps - 2 >= G) {
        ns_gpu_ctx::GraphKey other = graph_key(c, p, true);
        other.cur ^= 1;
        if (!same_key(graph_key(c, p, true), c->gkey) && same_key(other, c->gkey)) {
            rc = overlapped_step(c, p, c->ev);
            done++;
        }
    }
    const int64_t launches = graph ? (nsteps - done) / G : 0;
    bool joined = false;
    if (rc == NS_OK && launches > 0) {
        // the graph's first plan must follow this step's late kernel, an
        // event from outside the graph: join before the first launch
        hipError_t e = hipStreamWaitEvent(c->stream, c->ev[2], 0);
        joined = true;
        if (e != hipSuccess)
            rc = set_error(NS_ERR_HIP, "ns_gpu_run: join of the second stream: %s", hipGetErrorString(e));
        const ns_gpu_ctx::GraphKey key = graph_key(c, p, true);
        if (rc == NS_OK && (!c->graph_exec || !same_key(key, c->gkey)))
            rc = capture_overlapped(c, p, key, G);
        for (int64_t k = 0; k < launches && rc == NS_OK; k++) {
            if ((e = hipGraphLaunch(c->graph_exec, c->stream)) != hipSuccess)
                rc = set_error(NS_ERR_HIP, "ns_gpu_run: graph launch: %s", hipGetErrorString(e));
            done += G;
        }
    }
    for (; done < nsteps && rc == NS_OK; done++) {
        rc = overlapped_step(c, p, c->ev);
        joined = false;
    }
    if (!joined) {
        // join even after a failed step, so the second stream is never left
        // running past the call
        const hipError_t e = hipStreamWaitEvent(c->stream, c->ev[2], 0);
        if (rc == NS_OK && e != hipSuccess)
            rc = set_error(NS_ERR_HIP, "ns_gpu_run: join of the second stream: %s", hipGetErrorString(e));
    }
    if (rc != NS_OK) {
        c->keys_valid = false;  // late ranks may still hold the key sentinel
        return rc;
    }
    const void *r = c->f32 ? (const void *)c->rf[c->cur] : (const void *)c->r[c->cur];
    HIP_TRY(ns::overlap_fix_keys(c->ov, r, c->f32, c->sw.keys[0], c->n, c->keys_q24 ? c->qtab : nullptr,
                                 ns::quantile_stride(c->n), c->stream));
    return NS_OK;
}

} // namespace

extern "C" {

const char *ns_gpu_last_error(void) { return ns::g_err; }

const char *ns_gpu_version(void) { return "nsphere_amd 0.1 (gfx950, HIP)"; }

void ns_step_params_default(ns_step_params *p)
{
    if (!p)
        return;
    const double kmsec_to_kpcmyr = 1.02271e-3;
    p->dt = 1.0e-3;
    p->G = 4.3e-6;
    p->halo_mass = 1.0e12;
    p->vel_conv_sq = kmsec_to_kpcmyr * kmsec_to_kpcmyr;
    p->radius_tol = 1.0e-7;
    p->velocity_tol = 1.0e-7;
    p->alpha = 0.05;
    p->max_subdiv = 4096 * 4096 * 16;
    p->out_type = 2;
    p->N_taumin = 10;
    p->identity_gravity = 0;
}

void ns_step_params_unit(ns_step_params *p)
{
    if (!p)
        return;
    ns_step_params_default(p);
    p->G = 1.0;
    p->halo_mass = 1.0;
    p->vel_conv_sq = 1.0;
}

int ns_gpu_ctx_create(int64_t npts, int device, ns_gpu_ctx **out)
{
    ns::clear_error();
    if (!out)
        return set_error(NS_ERR_ARG, "ns_gpu_ctx_create: null out");
    *out = nullptr;
    if (npts < 0 || npts > (int64_t)INT32_MAX)
        return set_error(NS_ERR_ARG, "ns_gpu_ctx_create: npts %lld out of range [0, 2^31)", (long long)npts);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_error(NS_ERR_NODEV, "ns_gpu_ctx_create: no HIP device");
    if (device < 0 || device >= ndev)
        return set_error(NS_ERR_ARG, "ns_gpu_ctx_create: device %d not in [0,%d)", device, ndev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return set_error(NS_ERR_NODEV, "ns_gpu_ctx_create: cannot query device %d", device);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(NS_ERR_NODEV, "ns_gpu_ctx_create: device %d is %s, library built for gfx950", device,
                         prop.gcnArchName);

    DeviceGuard g(device);
    ns_gpu_ctx *c = new (std::nothrow) ns_gpu_ctx();
    if (!c)
        return set_error(NS_ERR_NOMEM, "ns_gpu_ctx_create: host allocation");
    c->device = device;
    c->n = npts;
    const int64_t n = npts;
    const int64_t tiles = ns::sort_tiles(n) > 0 ? ns::sort_tiles(n) : 1;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
        if (e == hipSuccess) e = dalloc(&c->r[b], n);
        if (e == hipSuccess) e = dalloc(&c->v[b], n);
        if (e == hipSuccess) e = dalloc(&c->l[b], n);
        if (e == hipSuccess) e = dalloc(&c->id[b], n);
        if (e == hipSuccess) e = dalloc(&c->sw.keys[b], n);
        if (e == hipSuccess) e = dalloc(&c->sw.vals[b], n);
    }
    if (e == hipSuccess) e = dalloc(&c->lvl, n);
    if (e == hipSuccess) e = dalloc(&c->perm, n);
    if (e == hipSuccess) e = dalloc(&c->inv, n);
    if (e == hipSuccess) e = dalloc(&c->sw.hist, 2 * ns::kKeyPasses * ns::kRadix);
    if (e == hipSuccess) e = dalloc(&c->sw.lookback, tiles * ns::kRadix);
    if (e == hipSuccess) e = dalloc(&c->sw.counters, 64);
    if (e == hipSuccess) e = dalloc(&c->energy_partial, 2 * ns::energy_blocks(n));
    if (e == hipSuccess) e = dalloc(&c->energy_out, 2);
    if (e == hipSuccess) e = dalloc(&c->budget_hits, 1);
    c->work_stride = ns::work_stride_for(npts);
    if (e == hipSuccess) e = dalloc(&c->work, ns::kWorkKernels * c->work_stride + ns::kChainWords);
    if (e == hipSuccess) e = dalloc(&c->q_lf, n);
    if (e == hipSuccess) e = dalloc(&c->q_lc, n);
    if (e == hipSuccess) e = dalloc(&c->q_lcw, n);
    if (e == hipSuccess) e = dalloc(&c->q_count, ns::kQueueCounters);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemset(c->sw.lookback, 0, (size_t)tiles * ns::kRadix * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemset(c->budget_hits, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->sw.counters, 0, 64 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(c->work, 0, work_bytes(c));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory) ? NS_ERR_NOMEM : NS_ERR_HIP;
        set_error(code, "ns_gpu_ctx_create(npts=%lld): %s", (long long)npts, hipGetErrorString(e));
        free_all(c);
        if (c->stream)
            (void)hipStreamDestroy(c->stream);
        delete c;
        return code;
    }
    c->sw.epoch = 1;
    c->prof.stream = c->stream;
    *out = c;
    return NS_OK;
}

int ns_gpu_ctx_destroy(ns_gpu_ctx *c)
{
    if (!c)
        return NS_OK;
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->side)
        (void)hipStreamSynchronize(c->side);
    c->prof.release();
    release_graph(c);
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NS_OK;
}

int64_t ns_gpu_npts(const ns_gpu_ctx *c) { return c ? c->n : -1; }

int ns_gpu_set_option(ns_gpu_ctx *c, int option, int64_t value)
{
    if (int rc = check_ctx(c, "ns_gpu_set_option"))
        return rc;
    release_graph(c);  // any option may change what a captured step does
    switch (option) {
    case NS_OPT_GRAPH:
        if (value != 0 && value != 1)
            return set_error(NS_ERR_ARG, "NS_OPT_GRAPH takes 0 or 1");
        c->use_graph = value == 1;
        return NS_OK;
    case NS_OPT_LANE_LEVELS:
        if (value < 1 || value > 6)
            return set_error(NS_ERR_ARG, "NS_OPT_LANE_LEVELS takes 1..6 (leapfrog levels 0..value-1 per lane)");
        c->cap_N = (int64_t)1 << (value - 1);
        return NS_OK;
    case NS_OPT_DEEP_BLOCKS:
        if (value < 1 || value > 4096)
            return set_error(NS_ERR_ARG, "NS_OPT_DEEP_BLOCKS takes 1..4096");
        c->deep_blocks = (int)value;
        return NS_OK;
    case NS_OPT_TIEBREAK:
        if (value != 0 && value != 1)
            return set_error(NS_ERR_ARG, "NS_OPT_TIEBREAK takes 0 or 1");
        if (value == 1 && !c->id_order) {
            DeviceGuard g(c->device);
            hipError_t e = dalloc(&c->id_order, c->n);
            if (e != hipSuccess)
                return set_error(NS_ERR_NOMEM, "NS_OPT_TIEBREAK: %s", hipGetErrorString(e));
        }
        if (c->tiebreak != (int)value)
            release_graph(c);
        c->tiebreak = (int)value;
        return NS_OK;
    case NS_OPT_PROFILE:
        c->prof.on = value != 0;
        return NS_OK;
    case NS_OPT_PROFILE_KERNELS:
        if (value < 0 || value > 0xFFFFFFFFll)
            return set_error(NS_ERR_ARG, "NS_OPT_PROFILE_KERNELS takes a 32-bit mask of kernel families");
        c->prof.mask = (uint32_t)value;
        return NS_OK;
    case NS_OPT_OVERLAP:
        if (value != 0 && value != 1)
            return set_error(NS_ERR_ARG, "NS_OPT_OVERLAP takes 0 or 1");
        c->overlap = (int)value;
        return NS_OK;
    case NS_OPT_GRAPH_STEPS:
        if (value < 2 || value > 64 || (value & 1))
            return set_error(NS_ERR_ARG, "NS_OPT_GRAPH_STEPS takes an even count in 2..64");
        c->graph_steps = (int)value;
        return NS_OK;
    case NS_OPT_SORT_KEYS:
        if (value < 0 || value > 2)
            return set_error(NS_ERR_ARG, "NS_OPT_SORT_KEYS takes 0, 1 or 2");
        c->sort_keys = (int)value;
        return NS_OK;
    case NS_OPT_INTEGRATOR:
        if (value < ns::FORM_PLAIN || value > ns::FORM_BINNED_FAST)
            return set_error(NS_ERR_ARG, "NS_OPT_INTEGRATOR takes 0 (plain), 1 (binned) or 2 (binned, fast)");
        c->integrator = (int)value;
        return NS_OK;
    case NS_OPT_BUDGET:
        if (value != 0 && value < 2)
            return set_error(NS_ERR_ARG, "NS_OPT_BUDGET must be 0 (no cap) or >= 2 force evaluations");
        c->budget = value == 0 ? INT64_MAX : value;
        return NS_OK;
    case NS_OPT_SORT_TILE: {
        ns::SortConfig cfg = c->sw.config;
        cfg.threads = (int)(value >> 16);
        cfg.items = (int)(value & 0xFFFF);
        if (!ns::sort_config_supported(cfg))
            return set_error(NS_ERR_ARG, "NS_OPT_SORT_TILE: (threads << 16 | items) with threads x items in "
                                         "{256x16, 256x24, 512x8, 512x12, 512x16, 1024x6}");
        c->sw.config = cfg;
        return NS_OK;
    }
    case NS_OPT_SORT_ALG: {
        ns::SortConfig cfg = c->sw.config;
        cfg.alg = (int)value;
        if (value < 0 || value > 1 || !ns::sort_config_supported(cfg))
            return set_error(NS_ERR_ARG, "NS_OPT_SORT_ALG takes 0 (reduce-then-scan) or 1 (onesweep)");
        if (value == ns::kSortOnesweep && c->mlp)
            return set_error(NS_ERR_STATE, "NS_OPT_SORT_ALG: learned binning sorts by reduce-then-scan only");
        if (value == ns::kSortOnesweep && c->nseg > 0)
            return set_error(NS_ERR_STATE, "NS_OPT_SORT_ALG: an ensemble sorts by segments, reduce-then-scan only");
        c->sw.config = cfg;
        return NS_OK;
    }
    case NS_OPT_PRECISION: {
        if (value != 0 && value != 1)
            return set_error(NS_ERR_ARG, "NS_OPT_PRECISION takes 0 (float64 state) or 1 (float32 state)");
        const bool want = value == 1;
        if (want == c->f32)
            return NS_OK;
        DeviceGuard g(c->device);
        if (want)
            if (int rc = ensure_f32(c))
                return rc;
        if (c->uploaded && c->n > 0) {  // carry the current state over
            if (int rc = want ? state_to_f32(c, c->cur) : state_to_f64(c, c->cur))
                return rc;
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
        c->f32 = want;
        if (c->has_mass && c->n > 0) {  // and the masses
            if (int rc = ensure_mass(c))
                return rc;
            if (want)
                HIP_TRY((ns::launch_convert<float, double>(c->md[c->cur], c->mf[c->cur], c->n, c->stream)));
            else
                HIP_TRY((ns::launch_convert<double, float>(c->mf[c->cur], c->md[c->cur], c->n, c->stream)));
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
        c->keys_valid = false;
        return NS_OK;
    }
    default:
        return set_error(NS_ERR_ARG, "ns_gpu_set_option: unknown option %d", option);
    }
}

} // extern "C"

namespace {

// An upload writes the state into buffer 0.  Per-particle masses stay by
// position (mass k goes with the particle uploaded at position k), so after
// an odd number of reorderings they move to buffer 0 with it.
int masses_to_buffer0(ns_gpu_ctx *c)
{
    if (!c->has_mass || c->cur == 0 || c->n == 0)
        return NS_OK;
    if (c->f32)
        HIP_TRY(hipMemcpyAsync(c->mf[0], c->mf[c->cur], (size_t)c->n * sizeof(float), hipMemcpyDeviceToDevice,
                               c->stream));
    else
        HIP_TRY(hipMemcpyAsync(c->md[0], c->md[c->cur], (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    return NS_OK;
}

} // namespace

extern "C" {

int ns_gpu_upload_soa(ns_gpu_ctx *c, const double *r, const double *v, const double *ell, const uint32_t *id)
{
    if (int rc = check_ctx(c, "ns_gpu_upload_soa"))
        return rc;
    if (c->n > 0 && (!r || !v || !ell || !id))
        return set_error(NS_ERR_ARG, "ns_gpu_upload_soa: null array");
    DeviceGuard g(c->device);
    const size_t nd = (size_t)c->n * sizeof(double), ni = (size_t)c->n * sizeof(uint32_t);
    if (int rc = masses_to_buffer0(c))
        return rc;
    c->cur = 0;
    c->last_perm = nullptr;
    if (c->n > 0) {
        HIP_TRY(hipMemcpyAsync(c->r[0], r, nd, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->v[0], v, nd, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->l[0], ell, nd, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->id[0], id, ni, hipMemcpyHostToDevice, c->stream));
        if (c->f32)
            if (int rc = state_to_f32(c, 0))
                return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->keys_valid = false;
    c->levels_valid = false;
    return NS_OK;
}

// The state just written into the float64 buffers of index 0: convert for
// the float32 form and mark it uploaded, as ns_gpu_upload_soa does.
static int state_written(ns_gpu_ctx *c)
{
    if (int rc = masses_to_buffer0(c))
        return rc;
    c->cur = 0;
    c->last_perm = nullptr;
    if (c->n > 0 && c->f32)
        if (int rc = state_to_f32(c, 0))
            return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->keys_valid = false;
    c->levels_valid = false;
    return NS_OK;
}

int ns_gpu_generate_ic(ns_gpu_ctx *c, int kind, uint64_t seed, double rmax, double lowl_frac, double lowl_scale)
{
    if (int rc = check_ctx(c, "ns_gpu_generate_ic"))
        return rc;
    if (kind < 0 || kind > 2 || !(rmax > 0.0) || c->n > (int64_t)UINT32_MAX)
        return set_error(NS_ERR_ARG, "ns_gpu_generate_ic: bad kind/rmax");
    DeviceGuard g(c->device);
    HIP_TRY(ns::launch_generate_ic(kind, seed, rmax, lowl_frac, lowl_scale, c->n, c->r[0], c->v[0], c->l[0], c->id[0],
                                   c->stream));
    return state_written(c);
}

int ns_gpu_upload_ic(ns_gpu_ctx *c, const double *records, double conv)
{
    if (int rc = check_ctx(c, "ns_gpu_upload_ic"))
        return rc;
    if (c->n > 0 && !records)
        return set_error(NS_ERR_ARG, "ns_gpu_upload_ic: null records");
    for (int64_t i = 0; i < c->n; i++) {
        const double x = records[5 * i + 3];
        if (!(x >= 0.0 && x < 4294967296.0))
            return set_error(NS_ERR_ARG, "ns_gpu_upload_ic: id %g at %lld is not a uint32", x, (long long)i);
    }
    if (c->n == 0)
        return state_written(c);
    DeviceGuard g(c->device);
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, (size_t)c->n * 5 * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, records, (size_t)c->n * 5 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = ns::launch_ic_records(d, c->n, conv, c->r[0], c->v[0], c->l[0], c->id[0], c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_upload_ic: %s", hipGetErrorString(e));
    return state_written(c);
}

int ns_gpu_upload(ns_gpu_ctx *c, double *const particles[5])
{
    if (int rc = check_ctx(c, "ns_gpu_upload"))
        return rc;
    if (!particles || (c->n > 0 && (!particles[0] || !particles[1] || !particles[2] || !particles[3])))
        return set_error(NS_ERR_ARG, "ns_gpu_upload: null particle column");
    uint32_t *ids = (uint32_t *)malloc((size_t)(c->n > 0 ? c->n : 1) * sizeof(uint32_t));
    if (!ids)
        return set_error(NS_ERR_NOMEM, "ns_gpu_upload: host allocation");
    for (int64_t i = 0; i < c->n; i++) {
        const double x = particles[3][i];
        if (!(x >= 0.0 && x < 4294967296.0)) {
            free(ids);
            return set_error(NS_ERR_ARG, "ns_gpu_upload: id %g at %lld is not a uint32", x, (long long)i);
        }
        ids[i] = (uint32_t)x;
    }
    int rc = ns_gpu_upload_soa(c, particles[0], particles[1], particles[2], ids);
    free(ids);
    return rc;
}

int ns_gpu_download_soa(ns_gpu_ctx *c, double *r, double *v, double *ell, uint32_t *id)
{
    if (int rc = check_ctx(c, "ns_gpu_download_soa"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_download_soa: no state uploaded");
    DeviceGuard g(c->device);
    const size_t nd = (size_t)c->n * sizeof(double), ni = (size_t)c->n * sizeof(uint32_t);
    const int s = c->cur;
    if (c->n > 0) {
        if (c->f32)  // widen into the float64 staging buffers of the same index
            if (int rc = state_to_f64(c, s))
                return rc;
        if (r) HIP_TRY(hipMemcpyAsync(r, c->r[s], nd, hipMemcpyDeviceToHost, c->stream));
        if (v) HIP_TRY(hipMemcpyAsync(v, c->v[s], nd, hipMemcpyDeviceToHost, c->stream));
        if (ell) HIP_TRY(hipMemcpyAsync(ell, c->l[s], nd, hipMemcpyDeviceToHost, c->stream));
        if (id) HIP_TRY(hipMemcpyAsync(id, c->id[s], ni, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int ns_gpu_download(ns_gpu_ctx *c, double *particles[5])
{
    if (int rc = check_ctx(c, "ns_gpu_download"))
        return rc;
    if (!particles || (c->n > 0 && (!particles[0] || !particles[1] || !particles[2] || !particles[3])))
        return set_error(NS_ERR_ARG, "ns_gpu_download: null particle column");
    uint32_t *ids = (uint32_t *)malloc((size_t)(c->n > 0 ? c->n : 1) * sizeof(uint32_t));
    if (!ids)
        return set_error(NS_ERR_NOMEM, "ns_gpu_download: host allocation");
    int rc = ns_gpu_download_soa(c, particles[0], particles[1], particles[2], ids);
    if (rc == NS_OK)
        for (int64_t i = 0; i < c->n; i++)
            particles[3][i] = (double)ids[i];
    free(ids);
    return rc;
}

int ns_gpu_device_state(ns_gpu_ctx *c, double **r, double **v, double **ell, uint32_t **id)
{
    if (int rc = check_ctx(c, "ns_gpu_device_state"))
        return rc;
    if (c->f32)
        return set_error(NS_ERR_STATE, "ns_gpu_device_state: the state is float32 (NS_OPT_PRECISION=1)");
    if (r) *r = c->r[c->cur];
    if (v) *v = c->v[c->cur];
    if (ell) *ell = c->l[c->cur];
    if (id) *id = c->id[c->cur];
    return NS_OK;
}

int ns_gpu_sort_particles(ns_gpu_ctx *c)
{
    if (int rc = check_ctx(c, "ns_gpu_sort_particles"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_sort_particles: no state uploaded");
    DeviceGuard g(c->device);
    if (int rc = sort_perm(c))
        return rc;
    if (int rc = mass_scan(c, c->perm, true))  // masses move with the particles
        return rc;
    const int s = c->cur, d = c->cur ^ 1;
    c->prof.begin(ns::K_GATHER);
    if (c->f32)
        HIP_TRY(ns::launch_gather_state<float>(c->rf[s], c->vf[s], c->lf[s], c->id[s], c->perm, c->rf[d], c->vf[d],
                                               c->lf[d], c->id[d], c->n, c->stream));
    else
        HIP_TRY(ns::launch_gather_state<double>(c->r[s], c->v[s], c->l[s], c->id[s], c->perm, c->r[d], c->v[d],
                                                c->l[d], c->id[d], c->n, c->stream));
    c->prof.end(ns::K_GATHER);
    c->last_perm = c->perm;
    c->cur = d;
    c->keys_valid = false;
    c->levels_valid = false;
    return NS_OK;
}

int ns_gpu_permutation(ns_gpu_ctx *c, uint32_t *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_permutation"))
        return rc;
    if (!host_out && c->n > 0)
        return set_error(NS_ERR_ARG, "ns_gpu_permutation: null output");
    if (!c->last_perm)
        return set_error(NS_ERR_STATE, "ns_gpu_permutation: no sort or step since the last upload");
    if (c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    // vals[0] is the sort's payload scratch, free between sorts
    HIP_TRY(ns::launch_perm_copy(c->last_perm, c->sw.vals[0], c->n, c->stream));
    HIP_TRY(hipMemcpyAsync(host_out, c->sw.vals[0], (size_t)c->n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int ns_gpu_permute_column(ns_gpu_ctx *c, const double *in, double *out)
{
    if (int rc = check_ctx(c, "ns_gpu_permute_column"))
        return rc;
    if ((!in || !out) && c->n > 0)
        return set_error(NS_ERR_ARG, "ns_gpu_permute_column: null column");
    if (!c->last_perm)
        return set_error(NS_ERR_STATE, "ns_gpu_permute_column: no sort or step since the last upload");
    if (c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    // staging: the key ping-pong buffer keys[1] (free between sorts) and the
    // float64 buffer cur^1, which holds the state before the reordering
    double *d_in = reinterpret_cast<double *>(c->sw.keys[1]);
    double *d_out = c->r[c->cur ^ 1];
    const size_t nd = (size_t)c->n * sizeof(double);
    HIP_TRY(hipMemcpyAsync(d_in, in, nd, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(ns::launch_gather_column(d_in, c->last_perm, d_out, c->n, c->stream));
    HIP_TRY(hipMemcpyAsync(out, d_out, nd, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int ns_gpu_inverse_map(ns_gpu_ctx *c, int32_t *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_inverse_map"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_inverse_map: no state uploaded");
    DeviceGuard g(c->device);
    c->prof.begin(ns::K_INVMAP);
    HIP_TRY(ns::launch_inverse_map(c->id[c->cur], c->inv, c->n, c->stream));
    c->prof.end(ns::K_INVMAP);
    if (host_out && c->n > 0) {
        HIP_TRY(hipMemcpyAsync(host_out, c->inv, (size_t)c->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return NS_OK;
}

int ns_gpu_snapshot_records(ns_gpu_ctx *c, ns_apd_record *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_snapshot_records"))
        return rc;
    if (!host_out)
        return set_error(NS_ERR_ARG, "ns_gpu_snapshot_records: null output");
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_snapshot_records: no state uploaded");
    if (c->n > (int64_t)INT32_MAX)
        return set_error(NS_ERR_ARG, "ns_gpu_snapshot_records: the record's rank is an int32");
    if (c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, (size_t)c->n * sizeof(ns_apd_record)));
    const int s = c->cur;
    c->prof.begin(ns::K_APD);
    hipError_t e = c->f32 ? ns::launch_apd_records(c->rf[s], c->vf[s], c->lf[s], true, c->id[s], c->n, d, c->stream)
                          : ns::launch_apd_records(c->r[s], c->v[s], c->l[s], false, c->id[s], c->n, d, c->stream);
    c->prof.end(ns::K_APD);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host_out, d, (size_t)c->n * sizeof(ns_apd_record), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_snapshot_records: %s", hipGetErrorString(e));
    return NS_OK;
}

int ns_gpu_potential(ns_gpu_ctx *c, const ns_step_params *p, double *psi, double *psi0)
{
    if (int rc = check_ctx(c, "ns_gpu_potential"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_potential"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_potential: no state uploaded");
    if (c->n == 0) {
        if (psi0)
            *psi0 = 0.0;
        return NS_OK;
    }
    DeviceGuard g(c->device);
    const double dm = p->halo_mass / (double)c->n;  // deltaM, nsphere.c:5179
    double *d_psi = nullptr, *work = nullptr;
    const int64_t nw = ns::potential_work_doubles(c->n);
    HIP_TRY(hipMalloc((void **)&work, (size_t)nw * sizeof(double)));
    hipError_t e = psi ? hipMalloc((void **)&d_psi, (size_t)c->n * sizeof(double)) : hipSuccess;
    double total = 0.0;
    if (e == hipSuccess) {
        const int s = c->cur;
        c->prof.begin(ns::K_POTENTIAL);
        // without psi only the tile and total passes run
        e = ns::launch_potential(c->f32 ? (const void *)c->rf[s] : (const void *)c->r[s], c->f32, c->n, dm, p->G,
                                 p->vel_conv_sq, d_psi, work, c->stream);
        c->prof.end(ns::K_POTENTIAL);
    }
    if (e == hipSuccess && psi)
        e = hipMemcpyAsync(psi, d_psi, (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&total, work + nw - 1, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (d_psi)
        (void)hipFree(d_psi);
    (void)hipFree(work);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_potential: %s", hipGetErrorString(e));
    if (psi0)
        *psi0 = p->G * dm * total * p->vel_conv_sq;  // PsiA(0), nsphere.c:7733-7738
    return NS_OK;
}

int ns_gpu_rank_files(ns_gpu_ctx *c, const ns_step_params *p, const char *unsorted_path, const char *sorted_path)
{
    if (int rc = check_ctx(c, "ns_gpu_rank_files"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_rank_files"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_rank_files: no state uploaded");
    if (c->nseg > 0)
        return set_error(NS_ERR_STATE, "ns_gpu_rank_files: one halo per context (the files rank a whole snapshot)");
    if (c->n > (int64_t)INT32_MAX)
        return set_error(NS_ERR_ARG, "ns_gpu_rank_files: the rows' rank is an int32");
    DeviceGuard g(c->device);
    const int s = c->cur;
    int io_failed = 0;
    c->prof.begin(ns::K_RANKFILES);
    const hipError_t e = ns::rank_files(
        c->sw, c->f32 ? (const void *)c->rf[s] : (const void *)c->r[s], c->f32 ? (const void *)c->vf[s] : (const void *)c->v[s],
        c->f32 ? (const void *)c->lf[s] : (const void *)c->l[s], c->f32, c->id[s], c->n, p->G, p->vel_conv_sq,
        p->halo_mass, unsorted_path, sorted_path, c->stream, &io_failed);
    c->prof.end(ns::K_RANKFILES);
    c->keys_valid = false;  // the radius sort of the records used the key buffers
    if (io_failed)
        return set_error(NS_ERR_IO, "ns_gpu_rank_files: cannot write %s / %s", unsorted_path ? unsorted_path : "-",
                         sorted_path ? sorted_path : "-");
    if (e == hipErrorInvalidValue)
        return set_error(NS_ERR_ARG, "ns_gpu_rank_files: the density estimate has fewer than three distinct radii");
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_rank_files: %s", hipGetErrorString(e));
    return NS_OK;
}

int ns_gpu_density(ns_gpu_ctx *c, const ns_step_params *p, double *density, ns_density_info *info)
{
    if (int rc = check_ctx(c, "ns_gpu_density"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_density"))
        return rc;
    if (!density && c->n > 0)
        return set_error(NS_ERR_ARG, "ns_gpu_density: null output");
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_density: no state uploaded");
    if (c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, (size_t)c->n * sizeof(double)));
    const int s = c->cur;
    c->prof.begin(ns::K_DENSITY);
    hipError_t e = ns::density_estimate(c->f32 ? (const void *)c->rf[s] : (const void *)c->r[s], c->f32, c->n,
                                        p->halo_mass, d, info, c->stream);
    c->prof.end(ns::K_DENSITY);
    if (e == hipErrorInvalidValue) {
        (void)hipFree(d);
        return set_error(NS_ERR_ARG, "ns_gpu_density: fewer than three distinct decimated radii or a degenerate "
                                     "range (the reference skips such a snapshot)");
    }
    if (e == hipSuccess)
        e = hipMemcpy(density, d, (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_density: %s", hipGetErrorString(e));
    return NS_OK;
}

int ns_gpu_gather_ids(ns_gpu_ctx *c, const uint32_t *ids, int k, double *r, double *v, double *ell, int32_t *rank)
{
    if (int rc = check_ctx(c, "ns_gpu_gather_ids"))
        return rc;
    if (k < 0 || (k > 0 && !ids))
        return set_error(NS_ERR_ARG, "ns_gpu_gather_ids: bad ids");
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_gather_ids: no state uploaded");
    if (k == 0 || c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    HIP_TRY(ns::launch_inverse_map(c->id[c->cur], c->inv, c->n, c->stream));
    uint32_t *d_ids = nullptr;
    double *d_out = nullptr;
    HIP_TRY(hipMalloc((void **)&d_ids, (size_t)k * sizeof(uint32_t)));
    hipError_t e = hipMalloc((void **)&d_out, (size_t)k * 4 * sizeof(double));
    std::vector<double> h((size_t)k * 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_ids, ids, (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
    const int s = c->cur;
    if (e == hipSuccess)
        e = c->f32 ? ns::launch_gather_ids(c->rf[s], c->vf[s], c->lf[s], true, c->inv, c->n, d_ids, k, d_out,
                                           c->stream)
                   : ns::launch_gather_ids(c->r[s], c->v[s], c->l[s], false, c->inv, c->n, d_ids, k, d_out,
                                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_ids);
    if (d_out)
        (void)hipFree(d_out);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_gather_ids: %s", hipGetErrorString(e));
    for (int j = 0; j < k; j++) {
        if (r)
            r[j] = h[4 * j];
        if (v)
            v[j] = h[4 * j + 1];
        if (ell)
            ell[j] = h[4 * j + 2];
        if (rank)
            rank[j] = (int32_t)h[4 * j + 3];
    }
    return NS_OK;
}

int ns_gpu_integrate(ns_gpu_ctx *c, const ns_step_params *p)
{
    if (int rc = check_ctx(c, "ns_gpu_integrate"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_integrate"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_integrate: no state uploaded");
    DeviceGuard g(c->device);
    return integrate(c, p, false);
}

int ns_gpu_step(ns_gpu_ctx *c, const ns_step_params *p)
{
    if (int rc = check_ctx(c, "ns_gpu_step"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_step"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_step: no state uploaded");
    DeviceGuard g(c->device);
    if (int rc = sort_perm(c))
        return rc;
    return integrate(c, p, true);
}

int ns_gpu_run(ns_gpu_ctx *c, const ns_step_params *p, int64_t nsteps)
{
    if (nsteps < 0)
        return set_error(NS_ERR_ARG, "ns_gpu_run: nsteps < 0");
    const bool graph = c && c->use_graph && !c->prof.on && c->sw.config.alg == ns::kSortReduceScan && nsteps >= 3;
    if (c && c->uploaded && overlap_applies(c, nsteps)) {
        if (int rc = check_params(p, "ns_gpu_run"))
            return rc;
        DeviceGuard g(c->device);
        return run_overlapped(c, p, nsteps, graph);
    }
    if (!graph) {
        for (int64_t s = 0; s < nsteps; s++)
            if (int rc = ns_gpu_step(c, p))
                return rc;
        return NS_OK;
    }
    // One step eagerly (radius keys from the integrate epilogue, segment
    // table built), or two when the graph already captured expects the other
    // buffer parity, then pairs of steps from a captured graph: two steps
    // return the double buffers to the captured parity.  Capture records
    // the launches without running them.
    if (int rc = ns_gpu_step(c, p))
        return rc;
    nsteps--;
    DeviceGuard g(c->device);
    if (c->graph_exec && nsteps >= 3) {
        ns_gpu_ctx::GraphKey other = graph_key(c, p, false);
        other.cur ^= 1;
        if (!same_key(graph_key(c, p, false), c->gkey) && same_key(other, c->gkey)) {
            if (int rc = ns_gpu_step(c, p))
                return rc;
            nsteps--;
        }
    }
    const ns_gpu_ctx::GraphKey key = graph_key(c, p, false);
    if (!c->graph_exec || !same_key(key, c->gkey)) {
        release_graph(c);
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = sort_perm(c);
        if (!rc)
            rc = integrate(c, p, true);
        if (!rc)
            rc = sort_perm(c);
        if (!rc)
            rc = integrate(c, p, true);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc || e != hipSuccess) {
            if (gr)
                (void)hipGraphDestroy(gr);
            if (rc)
                return rc;
            return set_error(NS_ERR_HIP, "ns_gpu_run: stream capture failed: %s", hipGetErrorString(e));
        }
        c->graph = gr;
        const hipError_t ei = hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            release_graph(c);
            return set_error(NS_ERR_HIP, "ns_gpu_run: graph instantiation failed: %s", hipGetErrorString(ei));
        }
        c->gkey = key;
    }
    for (int64_t k = 0; k < nsteps / 2; k++)
        HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
    if (nsteps & 1)
        return ns_gpu_step(c, p);
    return NS_OK;
}

int ns_gpu_graph_prepare(ns_gpu_ctx *c, const ns_step_params *p, int64_t nsteps)
{
    if (int rc = check_ctx(c, "ns_gpu_graph_prepare"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_graph_prepare"))
        return rc;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_graph_prepare: no state uploaded");
    const bool graph = c->use_graph && !c->prof.on && c->sw.config.alg == ns::kSortReduceScan;
    if (!graph || !overlap_applies(c, nsteps) || !whole_run_graph(nsteps))
        return NS_OK;  // that run captures nothing ahead
    DeviceGuard g(c->device);
    if (int rc = ensure_overlap(c))
        return rc;
    const ns_gpu_ctx::GraphKey key = graph_key(c, p, true, nsteps);
    if (c->graph_exec && same_key(key, c->gkey))
        return NS_OK;
    return capture_overlapped(c, p, key, nsteps);
}

int ns_gpu_sync(ns_gpu_ctx *c)
{
    if (int rc = check_ctx(c, "ns_gpu_sync"))
        return rc;
    DeviceGuard g(c->device);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->side)
        HIP_TRY(hipStreamSynchronize(c->side));
    return NS_OK;
}

int ns_gpu_cap_hits(ns_gpu_ctx *c, uint64_t *count, int reset)
{
    if (int rc = check_ctx(c, "ns_gpu_cap_hits"))
        return rc;
    if (!count)
        return set_error(NS_ERR_ARG, "ns_gpu_cap_hits: null count");
    DeviceGuard g(c->device);
    unsigned long long hits = 0;
    HIP_TRY(hipMemcpyAsync(&hits, c->budget_hits, sizeof hits, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *count = hits;
    if (reset) {
        HIP_TRY(hipMemsetAsync(c->budget_hits, 0, sizeof hits, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return NS_OK;
}

int ns_gpu_levels(ns_gpu_ctx *c, uint8_t *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_levels"))
        return rc;
    if (!host_out && c->n > 0)
        return set_error(NS_ERR_ARG, "ns_gpu_levels: null output");
    if (!c->levels_valid)
        return set_error(NS_ERR_STATE, "ns_gpu_levels: no integrate since the last upload/sort");
    DeviceGuard g(c->device);
    if (c->n > 0) {
        HIP_TRY(hipMemcpyAsync(host_out, c->lvl, (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return NS_OK;
}

int ns_gpu_energy(ns_gpu_ctx *c, const ns_step_params *p, double *K, double *W)
{
    if (int rc = check_ctx(c, "ns_gpu_energy"))
        return rc;
    if (!p || !K || !W)
        return set_error(NS_ERR_ARG, "ns_gpu_energy: null argument");
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_energy: no state uploaded");
    DeviceGuard g(c->device);
    if (c->n == 0) {
        *K = 0.0;
        *W = 0.0;
        return NS_OK;
    }
    // Ranks of the current radii (after a step the arrays hold the order of
    // that step's sort, not of the new radii).
    // (into a queue buffer, free between steps: c->perm keeps the last
    // reordering's permutation for ns_gpu_permutation)
    uint32_t *eperm = c->q_lf;
    if (int rc = sort_perm(c, eperm))
        return rc;
    if (int rc = mass_scan(c, eperm, false))
        return rc;
    const double m = p->halo_mass / (double)c->n;
    const double *menc = c->has_mass ? c->menc : nullptr;
    c->prof.begin(ns::K_ENERGY);
    const int s = c->cur;
    if (c->f32)
        HIP_TRY(ns::launch_energy<float>(c->rf[s], c->vf[s], c->lf[s], eperm, c->has_mass ? c->mf[s] : nullptr,
                                         menc, c->seg_off, c->seg_mass, c->nseg, c->n, m, p->vel_conv_sq * p->G,
                                         p->halo_mass, p->identity_gravity, c->energy_partial, c->energy_out,
                                         c->stream));
    else
        HIP_TRY(ns::launch_energy<double>(c->r[s], c->v[s], c->l[s], eperm, c->has_mass ? c->md[s] : nullptr,
                                          menc, c->seg_off, c->seg_mass, c->nseg, c->n, m, p->vel_conv_sq * p->G,
                                          p->halo_mass, p->identity_gravity, c->energy_partial, c->energy_out,
                                          c->stream));
    c->prof.end(ns::K_ENERGY);
    double out[2];
    HIP_TRY(hipMemcpyAsync(out, c->energy_out, sizeof out, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *K = out[0];
    *W = out[1];
    return NS_OK;
}

int ns_gpu_kernel_stats(ns_gpu_ctx *c, int *count, const char **names, int64_t *launches, double *total_ms)
{
    if (int rc = check_ctx(c, "ns_gpu_kernel_stats"))
        return rc;
    if (!count)
        return set_error(NS_ERR_ARG, "ns_gpu_kernel_stats: null count");
    static_assert(ns::K_COUNT <= NS_MAX_KERNEL_STATS, "kernel families exceed the documented array size");
    DeviceGuard g(c->device);
    HIP_TRY(c->prof.fold());
    *count = ns::K_COUNT;
    for (int k = 0; k < ns::K_COUNT; k++) {
        if (names) names[k] = ns::kernel_name(k);
        if (launches) launches[k] = c->prof.launches[k];
        if (total_ms) total_ms[k] = c->prof.ms[k];
    }
    return NS_OK;
}

int ns_gpu_kernel_stats_reset(ns_gpu_ctx *c)
{
    if (int rc = check_ctx(c, "ns_gpu_kernel_stats_reset"))
        return rc;
    DeviceGuard g(c->device);
    c->prof.reset();
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemset(c->work, 0, work_bytes(c)));
    return NS_OK;
}

namespace {

// Device copy of the current state buffer (and masses) for the autotuner.
struct Snapshot {
    void *r = nullptr, *v = nullptr, *l = nullptr, *m = nullptr;
    uint32_t *id = nullptr;
    size_t bytes = 0, mbytes = 0;
    unsigned long long hits = 0;
};

hipError_t snapshot_copy(ns_gpu_ctx *c, Snapshot &s, bool save)
{
    const int b = c->cur;
    void *r = c->f32 ? (void *)c->rf[b] : (void *)c->r[b];
    void *v = c->f32 ? (void *)c->vf[b] : (void *)c->v[b];
    void *l = c->f32 ? (void *)c->lf[b] : (void *)c->l[b];
    void *m = c->has_mass ? (c->f32 ? (void *)c->mf[b] : (void *)c->md[b]) : nullptr;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    hipError_t e;
    if ((e = hipMemcpyAsync(save ? s.r : r, save ? r : s.r, s.bytes, k, c->stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(save ? s.v : v, save ? v : s.v, s.bytes, k, c->stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(save ? s.l : l, save ? l : s.l, s.bytes, k, c->stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(save ? (void *)s.id : (void *)c->id[b], save ? (void *)c->id[b] : (void *)s.id,
                            (size_t)c->n * sizeof(uint32_t), k, c->stream)) != hipSuccess)
        return e;
    if (m && (e = hipMemcpyAsync(save ? s.m : m, save ? m : s.m, s.mbytes, k, c->stream)) != hipSuccess)
        return e;
    return hipStreamSynchronize(c->stream);
}

void snapshot_free(Snapshot &s)
{
    for (void *p : {s.r, s.v, s.l, s.m, (void *)s.id})
        if (p)
            (void)hipFree(p);
    s = Snapshot{};
}

} // namespace

int ns_gpu_autotune(ns_gpu_ctx *c, const ns_step_params *p, int steps, ns_tune_result *out)
{
    if (int rc = check_ctx(c, "ns_gpu_autotune"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_autotune"))
        return rc;
    if (!out || steps < 1 || steps > 10000)
        return set_error(NS_ERR_ARG, "ns_gpu_autotune: null result or steps outside 1..10000");
    if (!c->uploaded || c->n == 0)
        return set_error(NS_ERR_STATE, "ns_gpu_autotune: no state uploaded");
    DeviceGuard g(c->device);
    release_graph(c);
    Snapshot snap;
    const size_t el = c->f32 ? sizeof(float) : sizeof(double);
    snap.bytes = (size_t)c->n * el;
    snap.mbytes = c->has_mass ? snap.bytes : 0;
    hipError_t e = hipMalloc(&snap.r, snap.bytes);
    if (e == hipSuccess) e = hipMalloc(&snap.v, snap.bytes);
    if (e == hipSuccess) e = hipMalloc(&snap.l, snap.bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&snap.id, (size_t)c->n * sizeof(uint32_t));
    if (e == hipSuccess && c->has_mass) e = hipMalloc(&snap.m, snap.mbytes);
    if (e == hipSuccess) e = hipMemcpy(&snap.hits, c->budget_hits, sizeof snap.hits, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = snapshot_copy(c, snap, true);
    if (e != hipSuccess) {
        snapshot_free(snap);
        return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP, "ns_gpu_autotune: snapshot: %s",
                         hipGetErrorString(e));
    }
    const int cur0 = c->cur;
    const bool prof_on = c->prof.on;
    c->prof.on = false;
    hipEvent_t ea = nullptr, eb = nullptr;
    (void)hipEventCreate(&ea);
    (void)hipEventCreate(&eb);
    out->ntried = 0;
    auto current = [&]() {
        ns_tune_config t = {};
        t.sort_alg = c->sw.config.alg;
        t.sort_threads = c->sw.config.threads;
        t.sort_items = c->sw.config.items;
        int lv = 1;
        while (((int64_t)1 << (lv - 1)) < c->cap_N)
            lv++;
        t.lane_levels = lv;
        t.deep_blocks = c->deep_blocks;
        t.sort_keys = c->sort_keys;
        return t;
    };
    auto apply = [&](const ns_tune_config &t) {
        c->sw.config.alg = t.sort_alg;
        c->sw.config.threads = t.sort_threads;
        c->sw.config.items = t.sort_items;
        c->cap_N = (int64_t)1 << (t.lane_levels - 1);
        c->deep_blocks = t.deep_blocks;
        c->sort_keys = t.sort_keys;
    };
    int rc = NS_OK;
    // one trial: restore, one untimed step, `steps` timed steps
    auto trial = [&](ns_tune_config &t) -> int {
        apply(t);
        c->cur = cur0;
        c->keys_valid = false;
        if (snapshot_copy(c, snap, false) != hipSuccess)
            return set_error(NS_ERR_HIP, "ns_gpu_autotune: restore failed");
        if (int r2 = ns_gpu_step(c, p))
            return r2;
        HIP_TRY(hipEventRecord(ea, c->stream));
        for (int s = 0; s < steps; s++)
            if (int r2 = ns_gpu_step(c, p))
                return r2;
        HIP_TRY(hipEventRecord(eb, c->stream));
        HIP_TRY(hipEventSynchronize(eb));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ea, eb));
        t.ms_per_step = ms / steps;
        if (out->ntried < NS_TUNE_MAX)
            out->tried[out->ntried++] = t;
        return NS_OK;
    };
    const ns_tune_config orig = current();
    ns_tune_config best = orig;
    best.ms_per_step = 1e300;
    // knob 0, where quantile keys apply: the digits of the sort between
    // steps -- a 12-bit bucket digit and a 12-bit sub-bin (1), three 8-bit
    // digits (2), or the 64-bit keys' four 8-bit digits of their high half
    // and a fix-up (0)
    if (!c->f32 && c->nseg == 0 && c->n >= ns::kQMinN)
        for (int sk : {1, 2, 0}) {
            if (rc)
                break;
            ns_tune_config t = best;
            t.sort_keys = sk;
            if ((rc = trial(t)) == NS_OK && t.ms_per_step < best.ms_per_step)
                best = t;
        }
    // knob 1: pass algorithm x tile shape
    static const int kTiles[][2] = {{256, 16}, {256, 24}, {512, 8}, {512, 12}, {512, 16}, {1024, 6}};
    constexpr int kNumTiles = (int)(sizeof kTiles / sizeof kTiles[0]);
    for (int alg = 0; alg < 2 && !rc; alg++) {
        if (alg == ns::kSortOnesweep && c->nseg > 0)
            continue;  // an ensemble sorts by segments
        for (int k = 0; k < kNumTiles && !rc; k++) {
            ns_tune_config t = best;
            t.sort_alg = alg;
            t.sort_threads = kTiles[k][0];
            t.sort_items = kTiles[k][1];
            if ((rc = trial(t)) == NS_OK && t.ms_per_step < best.ms_per_step)
                best = t;
        }
    }
    // knob 2: leapfrog levels per lane; knob 3: the wave kernel's grid
    for (int lv = 2; lv <= 5 && !rc; lv++) {
        ns_tune_config t = best;
        t.lane_levels = lv;
        if ((rc = trial(t)) == NS_OK && t.ms_per_step < best.ms_per_step)
            best = t;
    }
    for (int blocks : {512, 1024, 2048, 4096}) {
        if (rc)
            break;
        ns_tune_config t = best;
        t.deep_blocks = blocks;
        if ((rc = trial(t)) == NS_OK && t.ms_per_step < best.ms_per_step)
            best = t;
    }
    // leave the state as it was, with the chosen options
    c->cur = cur0;
    c->last_perm = nullptr;
    c->keys_valid = false;
    c->levels_valid = false;
    const hipError_t er = snapshot_copy(c, snap, false);
    if (er == hipSuccess)
        (void)hipMemcpy(c->budget_hits, &snap.hits, sizeof snap.hits, hipMemcpyHostToDevice);
    snapshot_free(snap);
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    c->prof.on = prof_on;
    if (rc) {
        apply(orig);
        return rc;
    }
    if (er != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_autotune: restore failed: %s", hipGetErrorString(er));
    apply(best);
    out->best = best;
    return NS_OK;
}

int ns_gpu_set_masses(ns_gpu_ctx *c, const double *mass)
{
    if (int rc = check_ctx(c, "ns_gpu_set_masses"))
        return rc;
    release_graph(c);
    if (!mass) {
        c->has_mass = false;
        return NS_OK;
    }
    if (c->nseg > 0)
        return set_error(NS_ERR_STATE, "ns_gpu_set_masses: per-particle masses and halos do not combine");
    if (c->n == 0)
        return NS_OK;
    long double total = 0.0L;
    for (int64_t i = 0; i < c->n; i++) {
        if (!(mass[i] >= 0.0) || !std::isfinite(mass[i]))
            return set_error(NS_ERR_ARG, "ns_gpu_set_masses: mass[%lld] = %g is not a finite non-negative number",
                             (long long)i, mass[i]);
        total += mass[i];
    }
    if (!(total > 0.0L))
        return set_error(NS_ERR_ARG, "ns_gpu_set_masses: total mass is zero");
    DeviceGuard g(c->device);
    if (int rc = ensure_mass(c))
        return rc;
    // fixed-point scale of the scan's tile prefixes: a power of two with
    // total * scale < 2^62 (the float32 pairs' truncations stay inside)
    int ex = 0;
    (void)std::frexp((double)total, &ex);  // total < 2^ex
    c->ms.scale = std::ldexp(1.0, 61 - ex);
    HIP_TRY(hipMemcpy(c->md[c->cur], mass, (size_t)c->n * sizeof(double), hipMemcpyHostToDevice));
    if (c->f32) {
        HIP_TRY((ns::launch_convert<float, double>(c->md[c->cur], c->mf[c->cur], c->n, c->stream)));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    c->has_mass = true;
    return NS_OK;
}

int ns_gpu_set_halos(ns_gpu_ctx *c, int nhalos, const int64_t *counts, const double *halo_mass)
{
    if (int rc = check_ctx(c, "ns_gpu_set_halos"))
        return rc;
    release_graph(c);
    DeviceGuard g(c->device);
    if (nhalos == 0 || !counts) {
        c->nseg = 0;
        c->sw.seg.sizes.clear();
        return NS_OK;
    }
    if (nhalos < 0 || nhalos > (1 << 20) || !halo_mass)
        return set_error(NS_ERR_ARG, "ns_gpu_set_halos: nhalos %d out of range [1, 2^20] or no masses", nhalos);
    if (c->has_mass)
        return set_error(NS_ERR_STATE, "ns_gpu_set_halos: per-particle masses and halos do not combine");
    if (c->sw.config.alg != ns::kSortReduceScan)
        return set_error(NS_ERR_STATE, "ns_gpu_set_halos: needs the reduce-then-scan sort (NS_OPT_SORT_ALG 0)");
    std::vector<int64_t> off(nhalos + 1, 0);
    for (int h = 0; h < nhalos; h++) {
        if (counts[h] < 1)
            return set_error(NS_ERR_ARG, "ns_gpu_set_halos: halo %d has %lld particles", h, (long long)counts[h]);
        if (!(halo_mass[h] > 0.0) || !std::isfinite(halo_mass[h]))
            return set_error(NS_ERR_ARG, "ns_gpu_set_halos: halo %d mass %g", h, halo_mass[h]);
        off[h + 1] = off[h] + counts[h];
    }
    if (off[nhalos] != c->n)
        return set_error(NS_ERR_ARG, "ns_gpu_set_halos: counts add up to %lld, not npts %lld",
                         (long long)off[nhalos], (long long)c->n);
    // device halo tables and the sort's segment table (sized for the
    // smallest supported tile, kSortTile keys)
    ns::seg_table_release(c->sw.seg);
    if (c->seg_off)
        (void)hipFree(c->seg_off);
    if (c->seg_mass)
        (void)hipFree(c->seg_mass);
    c->seg_off = nullptr;
    c->seg_mass = nullptr;
    c->nseg = 0;
    ns::SegTable &t = c->sw.seg;
    t.sizes.assign(counts, counts + nhalos);
    t.capacity = ns::sort_tiles(c->n) + nhalos + 1;
    hipError_t e = dalloc(&c->seg_off, nhalos + 1);
    if (e == hipSuccess) e = dalloc(&c->seg_mass, nhalos);
    if (e == hipSuccess) e = dalloc(&t.tile_start, t.capacity + 1);
    if (e == hipSuccess) e = dalloc(&t.tile_seg, t.capacity);
    if (e == hipSuccess) e = dalloc(&t.seg_first, nhalos + 1);
    if (e == hipSuccess) e = dalloc(&t.seg_start, nhalos + 1);
    if (e == hipSuccess) e = dalloc(&t.base, (int64_t)nhalos * ns::kRadix);
    if (e == hipSuccess) e = dalloc(&t.counts, t.capacity * ns::kRadix);
    if (e == hipSuccess) e = hipMemcpy(c->seg_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->seg_mass, halo_mass, (size_t)nhalos * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        t.sizes.clear();
        return set_error(e == hipErrorOutOfMemory ? NS_ERR_NOMEM : NS_ERR_HIP, "ns_gpu_set_halos: %s",
                         hipGetErrorString(e));
    }
    c->nseg = nhalos;
    return NS_OK;
}

int ns_gpu_set_mlp(ns_gpu_ctx *c, const ns_mlp *mlp)
{
    if (int rc = check_ctx(c, "ns_gpu_set_mlp"))
        return rc;
    release_graph(c);
    if (!mlp) {
        c->mlp = nullptr;
        return NS_OK;
    }
    if (c->sw.config.alg != ns::kSortReduceScan)
        return set_error(NS_ERR_STATE, "ns_gpu_set_mlp: needs the reduce-then-scan sort (NS_OPT_SORT_ALG 0)");
    if (ns::mlp_device(mlp) != c->device)
        return set_error(NS_ERR_ARG, "ns_gpu_set_mlp: network on device %d, context on device %d",
                         ns::mlp_device(mlp), c->device);
    DeviceGuard g(c->device);
    if (!c->order) {
        hipError_t e = dalloc(&c->order, c->n);
        if (e != hipSuccess)
            return set_error(NS_ERR_NOMEM, "ns_gpu_set_mlp: %s", hipGetErrorString(e));
    }
    c->mlp = mlp;
    return NS_OK;
}

int ns_gpu_features(ns_gpu_ctx *c, const ns_step_params *p, float *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_features"))
        return rc;
    if (int rc = check_params(p, "ns_gpu_features"))
        return rc;
    if (!host_out)
        return set_error(NS_ERR_ARG, "ns_gpu_features: null output");
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_gpu_features: no state uploaded");
    if (c->n == 0)
        return NS_OK;
    DeviceGuard g(c->device);
    if (int rc = mass_scan(c, nullptr, false))
        return rc;
    ns::StepArgs a = make_args(c, p);
    const int s = c->cur;
    a.r0 = c->f32 ? (const void *)c->rf[s] : (const void *)c->r[s];
    a.v0 = c->f32 ? (const void *)c->vf[s] : (const void *)c->v[s];
    a.l0 = c->f32 ? (const void *)c->lf[s] : (const void *)c->l[s];
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, (size_t)c->n * 5 * sizeof(float)));
    hipError_t e = ns::launch_mlp_features(a, c->f32, d, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host_out, d, (size_t)c->n * 5 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_features: %s", hipGetErrorString(e));
    return NS_OK;
}

int ns_gpu_enclosed_mass(ns_gpu_ctx *c, double *host_out)
{
    if (int rc = check_ctx(c, "ns_gpu_enclosed_mass"))
        return rc;
    if (!host_out)
        return set_error(NS_ERR_ARG, "ns_gpu_enclosed_mass: null output");
    if (!c->has_mass)
        return set_error(NS_ERR_STATE, "ns_gpu_enclosed_mass: no per-particle masses set");
    DeviceGuard g(c->device);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(host_out, c->menc, (size_t)(c->n + 2) * sizeof(double), hipMemcpyDeviceToHost));
    return NS_OK;
}

int ns_gpu_chain_steps(ns_gpu_ctx *c, uint64_t *out)
{
    if (int rc = check_ctx(c, "ns_gpu_chain_steps"))
        return rc;
    if (!out)
        return set_error(NS_ERR_ARG, "ns_gpu_chain_steps: null output");
    DeviceGuard g(c->device);
    unsigned long long h[ns::kChainWords];
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->side)
        HIP_TRY(hipStreamSynchronize(c->side));
    HIP_TRY(hipMemcpy(h, c->work + ns::kWorkKernels * c->work_stride, sizeof(h), hipMemcpyDeviceToHost));
    *out = h[ns::kChainSum] + h[ns::kChainMax];  // the last launch is not folded yet
    return NS_OK;
}

int ns_gpu_chain_floor(int device, int64_t micro_steps, double *ns_per_step)
{
    ns::clear_error();
    if (!ns_per_step || micro_steps < 1000 || micro_steps > ((int64_t)1 << 34))
        return set_error(NS_ERR_ARG, "ns_gpu_chain_floor: null output or micro_steps outside [1e3, 2^34]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return set_error(NS_ERR_NODEV, "ns_gpu_chain_floor: no device %d", device);
    DeviceGuard g(device);
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    double *d = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipMalloc((void **)&d, sizeof(double));
    float ms = 0.f;
    if (e == hipSuccess) e = ns::launch_chain_floor(1000, d, s);  // warm the code object
    if (e == hipSuccess) e = hipEventRecord(a, s);
    if (e == hipSuccess) e = ns::launch_chain_floor(micro_steps, d, s);
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (d) (void)hipFree(d);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess)
        return set_error(NS_ERR_HIP, "ns_gpu_chain_floor: %s", hipGetErrorString(e));
    *ns_per_step = (double)ms * 1.0e6 / (double)(micro_steps & ~(int64_t)3);  // whole groups of four ran
    return NS_OK;
}

int ns_gpu_force_evals(ns_gpu_ctx *c, uint64_t out[3])
{
    if (int rc = check_ctx(c, "ns_gpu_force_evals"))
        return rc;
    if (!out)
        return set_error(NS_ERR_ARG, "ns_gpu_force_evals: null output");
    DeviceGuard g(c->device);
    const size_t count = (size_t)ns::kWorkKernels * (size_t)c->work_stride;
    unsigned long long *h = (unsigned long long *)malloc(count * sizeof(unsigned long long));
    if (!h)
        return set_error(NS_ERR_NOMEM, "ns_gpu_force_evals: host buffer");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
        e = hipMemcpy(h, c->work, count * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        free(h);
        return set_error(NS_ERR_HIP, "ns_gpu_force_evals: %s", hipGetErrorString(e));
    }
    static_assert(ns::kWorkKernels == 3, "ns_gpu_force_evals reports three rows");
    for (int k = 0; k < ns::kWorkKernels; k++) {
        uint64_t s = 0;
        for (int64_t j = 0; j < c->work_stride; j++)
            s += h[k * c->work_stride + j];
        out[k] = s;
    }
    free(h);
    return NS_OK;
}

} // extern "C"

// ------------------------------------------------------ snapshot writer

struct ns_apd_writer {
    ns_gpu_ctx *ctx = nullptr;  // used by writes only: closing needs no context
    int device = 0;
    FILE *f = nullptr;
    int block_size = 1;
    int64_t n = 0;
    long long start = 0;     // file size when opened (the writer appends)
    int64_t written = 0;     // snapshots appended to the file
    void *d_rec[2] = {};     // device records, one buffer per slot
    void *h_rec[2] = {};     // pinned host copies
    hipStream_t copy = nullptr;
    hipEvent_t made[2] = {}, copied[2] = {};
    bool in_flight[2] = {};  // slot copied (or copying) but not yet in the file
    int slot = 0;
};

namespace {

void apd_release(ns_apd_writer *w)
{
    for (int k = 0; k < 2; k++) {
        if (w->d_rec[k])
            (void)hipFree(w->d_rec[k]);
        if (w->h_rec[k])
            (void)hipHostFree(w->h_rec[k]);
        if (w->made[k])
            (void)hipEventDestroy(w->made[k]);
        if (w->copied[k])
            (void)hipEventDestroy(w->copied[k]);
    }
    if (w->copy)
        (void)hipStreamDestroy(w->copy);
    if (w->f)
        fclose(w->f);
    delete w;
}

// the snapshot of slot k: wait for its copy, then append it to the file
int apd_flush_slot(ns_apd_writer *w, int k)
{
    if (!w->in_flight[k])
        return NS_OK;
    HIP_TRY(hipEventSynchronize(w->copied[k]));
    const size_t bytes = (size_t)w->n * sizeof(ns_apd_record);
    if (fwrite(w->h_rec[k], 1, bytes, w->f) != bytes)
        return set_error(NS_ERR_IO, "ns_apd_writer: short write");
    w->in_flight[k] = false;
    w->written++;
    return NS_OK;
}

} // namespace

extern "C" {

int ns_apd_writer_open(ns_gpu_ctx *c, const char *path, int block_size, ns_apd_writer **out)
{
    if (int rc = check_ctx(c, "ns_apd_writer_open"))
        return rc;
    if (!out || !path || block_size < 1)
        return set_error(NS_ERR_ARG, "ns_apd_writer_open: bad argument");
    *out = nullptr;
    if (!c->uploaded)
        return set_error(NS_ERR_STATE, "ns_apd_writer_open: no state uploaded");
    if (c->n > (int64_t)INT32_MAX)
        return set_error(NS_ERR_ARG, "ns_apd_writer_open: the record's rank is an int32");
    ns_apd_writer *w = new (std::nothrow) ns_apd_writer;
    if (!w)
        return set_error(NS_ERR_NOMEM, "ns_apd_writer_open");
    w->ctx = c;
    w->device = c->device;
    w->block_size = block_size;
    w->n = c->n;
    w->f = fopen(path, "ab");
    if (!w->f) {
        apd_release(w);
        return set_error(NS_ERR_IO, "ns_apd_writer_open: cannot open %s for appending", path);
    }
    fseek(w->f, 0, SEEK_END);
    w->start = ftell(w->f);
    DeviceGuard g(c->device);
    const size_t bytes = (size_t)(w->n > 0 ? w->n : 1) * sizeof(ns_apd_record);
    hipError_t e = hipStreamCreateWithFlags(&w->copy, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = hipMalloc(&w->d_rec[k], bytes);
        if (e == hipSuccess)
            e = hipHostMalloc(&w->h_rec[k], bytes, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&w->made[k], hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&w->copied[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        apd_release(w);
        return set_error(NS_ERR_NOMEM, "ns_apd_writer_open: %s", hipGetErrorString(e));
    }
    *out = w;
    return NS_OK;
}

int ns_apd_writer_write(ns_apd_writer *w)
{
    if (!w)
        return set_error(NS_ERR_ARG, "ns_apd_writer_write: null writer");
    ns_gpu_ctx *c = w->ctx;
    if (c->n != w->n)
        return set_error(NS_ERR_STATE, "ns_apd_writer_write: context size changed");
    DeviceGuard g(c->device);
    const int k = w->slot;
    // the slot's previous snapshot must be in the file before its buffers are reused
    if (int rc = apd_flush_slot(w, k))
        return rc;
    if (w->n > 0) {
        const int s = c->cur;
        c->prof.begin(ns::K_APD);
        HIP_TRY(c->f32 ? ns::launch_apd_records(c->rf[s], c->vf[s], c->lf[s], true, c->id[s], c->n, w->d_rec[k],
                                                c->stream)
                       : ns::launch_apd_records(c->r[s], c->v[s], c->l[s], false, c->id[s], c->n, w->d_rec[k],
                                                c->stream));
        c->prof.end(ns::K_APD);
        HIP_TRY(hipEventRecord(w->made[k], c->stream));
        HIP_TRY(hipStreamWaitEvent(w->copy, w->made[k], 0));
        HIP_TRY(hipMemcpyAsync(w->h_rec[k], w->d_rec[k], (size_t)w->n * sizeof(ns_apd_record), hipMemcpyDeviceToHost,
                               w->copy));
        HIP_TRY(hipEventRecord(w->copied[k], w->copy));
    }
    w->in_flight[k] = true;
    w->slot = k ^ 1;
    // the other slot's snapshot (the previous call's) goes to the file now,
    // while this one is being made and copied
    return apd_flush_slot(w, k ^ 1);
}

int ns_apd_writer_close(ns_apd_writer *w, int64_t *kept)
{
    if (!w)
        return set_error(NS_ERR_ARG, "ns_apd_writer_close: null writer");
    int rc = NS_OK;
    {
        DeviceGuard g(w->device);
        // oldest first: the slot not written last
        const int first = w->slot;
        rc = apd_flush_slot(w, first);
        if (rc == NS_OK)
            rc = apd_flush_slot(w, first ^ 1);
    }
    const int64_t whole = (w->written / w->block_size) * w->block_size;
    if (rc == NS_OK) {
        // the reference appends only full blocks: drop a partial last block
        const long long size = w->start + (long long)whole * w->n * (long long)sizeof(ns_apd_record);
        if (fflush(w->f) != 0 || (whole < w->written && ftruncate(fileno(w->f), (off_t)size) != 0))
            rc = set_error(NS_ERR_IO, "ns_apd_writer_close: cannot finish the file");
    }
    if (kept)
        *kept = rc == NS_OK ? whole : 0;
    apd_release(w);
    return rc;
}

} // extern "C"
