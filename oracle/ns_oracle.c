/*
 * ns_oracle.c -- CPU restatement of NSphere's per-step evolution loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see ns_oracle.h).
 * Each function follows the cited lines of src/c/nsphere.c (NSphere v0.1.2)
 * literally: same expression order, same branch structure, same quirks
 * (rank-0 feels no gravity, rank frozen for the whole step, 4F-3C output,
 * the LC convergence metric divided by rho_F^2, linear overshoot
 * interpolation).  Written from the reference's behaviour, not copied.
 */
#include "ns_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ forces */

/* nsphere.c:502-515: -(VEL_CONV_SQ*G) * (rank/npts) * M / (r*r), or 0 under
 * identity gravity (nsphere.c:476). */
double orc_gravitational_force(double r, int rank, int npts, const orc_params *p)
{
    if (p->identity_gravity)
        return 0.0;
    if (p->menc)
        return -(p->vel_conv_sq * p->G) * p->menc[rank] / (r * r);
    return -(p->vel_conv_sq * p->G) * ((double)rank / (double)npts) * p->halo_mass / (r * r);
}

/* nsphere.c:532-535: ell^2 / r^3 */
double orc_effective_angular_force(double r, double ell)
{
    return (ell * ell) / (r * r * r);
}

/* nsphere.c:563-575: same prefactor over rho^2 */
double orc_gravitational_force_rho_v(double rho, int rank, int npts, const orc_params *p)
{
    if (p->identity_gravity)
        return 0.0;
    if (p->menc)
        return -(p->vel_conv_sq * p->G) * p->menc[rank] / (rho * rho);
    return -(p->vel_conv_sq * p->G) * ((double)rank / (double)npts) * p->halo_mass / (rho * rho);
}

/* nsphere.c:596-599: ell^2 / rho^4 */
double orc_effective_angular_force_rho_v(double rho, double ell)
{
    return (ell * ell) / (rho * rho * rho * rho);
}

/* nsphere.c:1777-1787 */
static inline double force_lc(int i, int npts, const orc_params *p, double ell, double rho)
{
    double grav_part = orc_gravitational_force_rho_v(rho, i, npts, p);
    double ang_part = orc_effective_angular_force_rho_v(rho, ell);
    return grav_part + ang_part;
}

/* ------------------------------------------------------- adaptive leapfrog */

/* nsphere.c:1550-1607: KDK with (subSteps-1)/2 drifts of midStep. */
void orc_micro_leapfrog(int i, int npts, double r_in, double v_in, double ell,
                        double h, int N, int subSteps, const orc_params *p,
                        double *r_out, double *v_out)
{
    double r_curr = r_in;
    double v_curr = v_in;
    double half_kick, mid_step;
    if (subSteps == (2 * N + 1)) {
        half_kick = h / (2.0 * N);
        mid_step = h / (1.0 * N);
    } else {
        half_kick = h / (4.0 * N);
        mid_step = h / (2.0 * N);
    }
    double force = orc_gravitational_force(r_curr, i, npts, p);
    double dvdt = force + orc_effective_angular_force(r_curr, ell);
    v_curr += half_kick * dvdt;

    int pairs = (subSteps - 1) / 2;
    for (int pp = 1; pp <= (pairs - 1); pp++) {
        r_curr += mid_step * v_curr;
        force = orc_gravitational_force(r_curr, i, npts, p);
        dvdt = force + orc_effective_angular_force(r_curr, ell);
        v_curr += mid_step * dvdt;
    }
    r_curr += mid_step * v_curr;
    force = orc_gravitational_force(r_curr, i, npts, p);
    dvdt = force + orc_effective_angular_force(r_curr, ell);
    v_curr += half_kick * dvdt;

    *r_out = r_curr;
    *v_out = v_curr;
}

static inline int level_of(int64_t N)
{
    int k = 0;
    while (((int64_t)1 << k) < N)
        k++;
    return k > ORC_LVL_MASK ? ORC_LVL_MASK : k;
}

/* Work cap (not in the reference; budget 0 = unlimited). */
typedef struct {
    int64_t left;
    int limited, hit;
} work_cap;

static inline work_cap cap_of(const orc_params *p)
{
    work_cap w = {p->budget, p->budget > 0, 0};
    return w;
}

static inline int cap_take(work_cap *w, int64_t n)
{
    if (!w->limited)
        return 1;
    if (n > w->left) {
        w->hit = 1;
        return 0;
    }
    w->left -= n;
    return 1;
}

static int adaptive_leap_capped(int i, int npts, double r_in, double v_in, double ell, double h,
                                const orc_params *p, double *r_out, double *v_out, work_cap *w);

/* nsphere.c:1648-1735: N = 1, 2, 4, ... <= max_subdiv; converge when both
 * relative differences are below tolerance; output per out_type. */
int orc_adaptive_full_leap(int i, int npts, double r_in, double v_in, double ell,
                           double h, const orc_params *p, double *r_out, double *v_out)
{
    work_cap w = cap_of(p);
    return adaptive_leap_capped(i, npts, r_in, v_in, ell, h, p, r_out, v_out, &w);
}

static int adaptive_leap_capped(int i, int npts, double r_in, double v_in, double ell, double h,
                                const orc_params *p, double *r_out, double *v_out, work_cap *w)
{
    int N = 1;
    double r_coarse = r_in, v_coarse = v_in;
    double r_fine = r_in, v_fine = v_in;
    int last_level = 0;

    while (N <= p->max_subdiv) {
        if (N == 1 && !cap_take(w, 2)) {
            *r_out = r_in;
            *v_out = v_in;
            return ORC_LVL_NOCONV | ORC_LVL_MASK;
        }
        orc_micro_leapfrog(i, npts, r_in, v_in, ell, h, N, (2 * N + 1), p, &r_coarse, &v_coarse);
        if (!cap_take(w, 2 * (int64_t)N + 1)) {
            /* coarse(N) equals the previous level's fine pass bit for bit */
            if (p->out_type == 0) {
                *r_out = r_coarse;
                *v_out = v_coarse;
            } else if (p->out_type == 1) {
                *r_out = r_fine;
                *v_out = v_fine;
            } else {
                *r_out = 4.0 * r_fine - 3.0 * r_coarse;
                *v_out = 4.0 * v_fine - 3.0 * v_coarse;
            }
            return last_level | ORC_LVL_NOCONV;
        }
        orc_micro_leapfrog(i, npts, r_in, v_in, ell, h, N, (4 * N + 1), p, &r_fine, &v_fine);
        last_level = level_of(N);

        double radius_diff = fabs(r_fine - r_coarse) / (fabs(r_fine) + 1.0e-30);
        double velocity_diff = fabs(v_fine - v_coarse) / (fabs(v_fine) + 1.0e-30);

        if ((radius_diff < p->radius_tol) && (velocity_diff < p->velocity_tol)) {
            if (p->out_type == 0) {
                *r_out = r_coarse;
                *v_out = v_coarse;
            } else if (p->out_type == 1) {
                *r_out = r_fine;
                *v_out = v_fine;
            } else {
                *r_out = 4.0 * r_fine - 3.0 * r_coarse;
                *v_out = 4.0 * v_fine - 3.0 * v_coarse;
            }
            return last_level;
        }
        N *= 2;
    }
    if (p->out_type == 0) {
        *r_out = r_coarse;
        *v_out = v_coarse;
    } else if (p->out_type == 1) {
        *r_out = r_fine;
        *v_out = v_fine;
    } else {
        *r_out = 4.0 * r_fine - 3.0 * r_coarse;
        *v_out = 4.0 * v_fine - 3.0 * v_coarse;
    }
    return last_level | ORC_LVL_NOCONV;
}

/* ------------------------------------------------------------ Levi-Civita */

/* nsphere.c:1990-2036: subSteps KDK steps in tau of dtau = h_tau/subSteps;
 * t accumulates dtau * rho_mid^2.  subSteps is 64-bit here: the reference's
 * int overflows only in its non-convergence fallback at N > 2^28. */
void orc_micro_lc(int i, int npts, double rho_in, double v_in, double t_in,
                  int64_t subSteps, double h_tau, const orc_params *p, double ell,
                  double *rho_out, double *v_out, double *t_out)
{
    double dtau = h_tau / (double)subSteps;
    double rho_curr = rho_in;
    double v_curr = v_in;
    double t_curr = t_in;
    for (int64_t ss = 0; ss < subSteps; ss++) {
        double f1 = force_lc(i, npts, p, ell, rho_curr);
        double v_half = v_curr + 0.5 * dtau * f1;
        double rho_next = rho_curr + dtau * (0.5 * rho_curr * v_half);
        double f2 = force_lc(i, npts, p, ell, rho_next);
        double v_next = v_half + 0.5 * dtau * f2;
        double rho_mid = 0.5 * (rho_curr + rho_next);
        double dt_phys = dtau * (rho_mid * rho_mid);
        t_curr += dt_phys;
        rho_curr = rho_next;
        v_curr = v_next;
    }
    *rho_out = rho_curr;
    *v_out = v_curr;
    *t_out = t_curr;
}

/* nsphere.c:2077-2186: coarse 2N+1 vs fine 4N+1 substeps over the same h;
 * metric |rhoF-rhoC| / (rhoF^2 + 1e-30); 4F-3C on rho, v and t with rho
 * floored at 0; on non-convergence the fine pass at the overflowed N. */
static int single_tau_capped(int i, int npts, double rho_in, double v_in, double t_in,
                             double h_guess, const orc_params *p, double ell, double *rho_out,
                             double *v_out, double *t_out, work_cap *w);

int orc_single_tau_step_lc(int i, int npts, double rho_in, double v_in, double t_in,
                           double h_guess, const orc_params *p, double ell,
                           double *rho_out, double *v_out, double *t_out)
{
    work_cap w = cap_of(p);
    return single_tau_capped(i, npts, rho_in, v_in, t_in, h_guess, p, ell, rho_out, v_out, t_out,
                             &w);
}

static int single_tau_capped(int i, int npts, double rho_in, double v_in, double t_in,
                             double h_guess, const orc_params *p, double ell, double *rho_out,
                             double *v_out, double *t_out, work_cap *w)
{
    int64_t N = 1;
    int last_level = 0;
    double rhoF = 0.0, vF = 0.0, tF = 0.0; /* last fine pass, kept for the capped fallback */
    while (N <= p->max_subdiv) {
        double rhoC, vC, tC;
        if (N == 1 && !cap_take(w, 2 * N + 2)) {
            *rho_out = rho_in;
            *v_out = v_in;
            *t_out = t_in;
            return ORC_LVL_NOCONV | ORC_LVL_MASK;
        }
        orc_micro_lc(i, npts, rho_in, v_in, t_in, 2 * N + 1, h_guess, p, ell, &rhoC, &vC, &tC);
        if (!cap_take(w, 4 * N + 2)) {
            /* coarse(N) equals the previous level's fine pass bit for bit */
            *rho_out = rhoC;
            *v_out = vC;
            *t_out = tC;
            return last_level | ORC_LVL_NOCONV;
        }
        orc_micro_lc(i, npts, rho_in, v_in, t_in, 4 * N + 1, h_guess, p, ell, &rhoF, &vF, &tF);
        last_level = level_of(N);

        double rF = rhoF * rhoF;
        double rhodif = fabs(rhoF - rhoC) / (fabs(rF) + 1.0e-30);
        double vdif = fabs(vF - vC) / (fabs(vF) + 1.0e-30);

        if ((rhodif < p->radius_tol) && (vdif < p->velocity_tol)) {
            if (p->out_type == 0) {
                *rho_out = rhoC;
                *v_out = vC;
                *t_out = tC;
            } else if (p->out_type == 1) {
                *rho_out = rhoF;
                *v_out = vF;
                *t_out = tF;
            } else {
                double rho_rich = 4.0 * rhoF - 3.0 * rhoC;
                double v_rich = 4.0 * vF - 3.0 * vC;
                double t_rich = 4.0 * tF - 3.0 * tC;
                if (rho_rich < 0.0)
                    rho_rich = 0.0;
                *rho_out = rho_rich;
                *v_out = v_rich;
                *t_out = t_rich;
            }
            return last_level;
        }
        N *= 2;
    }
    if (!cap_take(w, 4 * N + 2)) {
        *rho_out = rhoF;
        *v_out = vF;
        *t_out = tF;
        return last_level | ORC_LVL_NOCONV;
    }
    orc_micro_lc(i, npts, rho_in, v_in, t_in, 4 * N + 1, h_guess, p, ell, rho_out, v_out, t_out);
    return level_of(N) | ORC_LVL_NOCONV;
}

/* nsphere.c:2231-2334: rho = sqrt(r); dtau = dt / (2 r N_taumin); tau steps
 * until t >= dt with linear interpolation on overshoot; stepMax = 1e8. */
int orc_adaptive_full_lc(int i, int npts, double r_in, double v_in, double ell,
                         double dt, const orc_params *p, double *r_out, double *v_out,
                         int64_t *n_tau_steps)
{
    int rec = ORC_LVL_LC;
    if (n_tau_steps)
        *n_tau_steps = 0;
    if (r_in < 1.0e-30) {
        *r_out = r_in;
        *v_out = v_in;
        return rec;
    }
    double rho_current = sqrt(r_in);
    double v_current = v_in;
    double t_cur = 0.0;
    double delta_tau;
    if (r_in > 1.0e-30 && p->N_taumin > 0)
        delta_tau = dt / (2.0 * r_in * p->N_taumin);
    else
        delta_tau = dt / 100.0;

    int64_t step_count = 0;
    const int64_t step_max = 100000000;
    int64_t taken = 0;
    work_cap w = cap_of(p); /* one cap for all tau steps of this particle-step */
    while (1) {
        if (t_cur >= dt)
            break;
        double rho_next, v_next, t_next;
        int lv = single_tau_capped(i, npts, rho_current, v_current, t_cur, delta_tau, p, ell,
                                   &rho_next, &v_next, &t_next, &w);
        taken++;
        if ((lv & ORC_LVL_MASK) > (rec & ORC_LVL_MASK))
            rec = (rec & ~ORC_LVL_MASK) | (lv & ORC_LVL_MASK);
        rec |= (lv & ORC_LVL_NOCONV);
        if (t_next > dt) {
            double alpha = 0.0;
            if (fabs(t_next - t_cur) > 1.0e-30)
                alpha = (dt - t_cur) / (t_next - t_cur);
            double rho_final = rho_current + alpha * (rho_next - rho_current);
            double v_final = v_current + alpha * (v_next - v_current);
            *r_out = rho_final * rho_final;
            *v_out = v_final;
            if (n_tau_steps)
                *n_tau_steps = taken;
            return rec;
        }
        rho_current = rho_next;
        v_current = v_next;
        t_cur = t_next;
        step_count++;
        if (step_count > step_max || w.hit) {
            *r_out = rho_current * rho_current;
            *v_out = v_current;
            if (n_tau_steps)
                *n_tau_steps = taken;
            return rec | ORC_LVL_NOCONV;
        }
    }
    *r_out = rho_current * rho_current;
    *v_out = v_current;
    if (n_tau_steps)
        *n_tau_steps = taken;
    return rec;
}

/* ------------------------------------------------------------------- sort */

uint64_t orc_radius_key(double r)
{
    uint64_t u;
    if (r != r) {
        u = 0x7FF8000000000000ull; /* one NaN, after +inf */
    } else if (r == 0.0) {
        u = 0; /* -0 and +0 compare equal under compare_particles */
    } else {
        memcpy(&u, &r, sizeof u);
    }
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

typedef struct {
    uint64_t key;
    uint32_t tie; /* id (tiebreak_id) or 0 */
    uint32_t idx;
} sort_item;

static inline int item_less(const sort_item *a, const sort_item *b)
{
    if (a->key != b->key)
        return a->key < b->key;
    return a->tie < b->tie;
}

/* Stable merge of [a, a+na) and [b, b+nb) into out: ties take the left run. */
static void merge_runs(const sort_item *a, int64_t na, const sort_item *b, int64_t nb,
                       sort_item *out)
{
    int64_t i = 0, j = 0, k = 0;
    while (i < na && j < nb) {
        if (item_less(&b[j], &a[i]))
            out[k++] = b[j++];
        else
            out[k++] = a[i++];
    }
    while (i < na)
        out[k++] = a[i++];
    while (j < nb)
        out[k++] = b[j++];
}

/* Stable bottom-up merge sort of items[0..n); tmp has n slots. */
static void stable_sort_items(sort_item *items, sort_item *tmp, int64_t n)
{
    /* insertion sort runs of 32 (stable) */
    const int64_t run = 32;
    for (int64_t s = 0; s < n; s += run) {
        int64_t e = s + run < n ? s + run : n;
        for (int64_t i = s + 1; i < e; i++) {
            sort_item x = items[i];
            int64_t j = i - 1;
            while (j >= s && item_less(&x, &items[j])) {
                items[j + 1] = items[j];
                j--;
            }
            items[j + 1] = x;
        }
    }
    sort_item *src = items, *dst = tmp;
    for (int64_t w = run; w < n; w *= 2) {
        for (int64_t s = 0; s < n; s += 2 * w) {
            int64_t m = s + w < n ? s + w : n;
            int64_t e = s + 2 * w < n ? s + 2 * w : n;
            merge_runs(src + s, m - s, src + m, e - m, dst + s);
        }
        sort_item *t = src;
        src = dst;
        dst = t;
    }
    if (src != items)
        memcpy(items, src, (size_t)n * sizeof(sort_item));
}

/* Parallel stable sort: stable-sort P contiguous chunks, then merge pairs in
 * rounds (left run first on ties), so the result equals a sequential stable sort. */
static void parallel_stable_sort(sort_item *items, int64_t n)
{
    sort_item *tmp = (sort_item *)malloc((size_t)(n > 0 ? n : 1) * sizeof(sort_item));
    int P = 1;
#ifdef _OPENMP
    P = omp_get_max_threads();
#endif
    if (n < 65536)
        P = 1;
    int chunks = 1;
    while (chunks < P)
        chunks *= 2;
    int64_t csz = (n + chunks - 1) / chunks;
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < chunks; c++) {
        int64_t s = (int64_t)c * csz;
        int64_t e = s + csz < n ? s + csz : n;
        if (s < e)
            stable_sort_items(items + s, tmp + s, e - s);
    }
    sort_item *src = items, *dst = tmp;
    for (int64_t w = csz; w < n; w *= 2) {
        int64_t npairs = (n + 2 * w - 1) / (2 * w);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t q = 0; q < npairs; q++) {
            int64_t s = q * 2 * w;
            int64_t m = s + w < n ? s + w : n;
            int64_t e = s + 2 * w < n ? s + 2 * w : n;
            merge_runs(src + s, m - s, src + m, e - m, dst + s);
        }
        sort_item *t = src;
        src = dst;
        dst = t;
    }
    if (src != items)
        memcpy(items, src, (size_t)n * sizeof(sort_item));
    free(tmp);
}

void orc_sort_perm(const double *r, const uint32_t *id, int64_t n, int tiebreak_id,
                   uint32_t *perm)
{
    sort_item *items = (sort_item *)malloc((size_t)(n > 0 ? n : 1) * sizeof(sort_item));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        items[i].key = orc_radius_key(r[i]);
        items[i].tie = (tiebreak_id && id) ? id[i] : 0u;
        items[i].idx = (uint32_t)i;
    }
    parallel_stable_sort(items, n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        perm[i] = items[i].idx;
    free(items);
}

/* nsphere.c:8475-8542 with the constants of nsphere.c:8280-8282: 24 chunks
 * sorted independently, then windows of max(minChunk/8, 50) around each seam
 * re-sorted.  Chunk and window sorts are stable, as quadsort is. */
void orc_sort_perm_chunked(const double *r, int64_t n, uint32_t *perm)
{
    const int sections = 24, overlap_div = 8, min_overlap = 50;
    sort_item *items = (sort_item *)malloc((size_t)(n > 0 ? n : 1) * sizeof(sort_item));
    sort_item *tmp = (sort_item *)malloc((size_t)(n > 0 ? n : 1) * sizeof(sort_item));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        items[i].key = orc_radius_key(r[i]);
        items[i].tie = 0;
        items[i].idx = (uint32_t)i;
    }
    if (n < sections) {
        stable_sort_items(items, tmp, n);
    } else {
        int64_t start[24], end[24];
        int64_t base = n / sections, rem = n % sections, off = 0;
        for (int c = 0; c < sections; c++) {
            int64_t sz = base + (c < rem ? 1 : 0);
            start[c] = off;
            end[c] = off + sz - 1;
            off += sz;
        }
#pragma omp parallel for schedule(dynamic)
        for (int c = 0; c < sections; c++)
            stable_sort_items(items + start[c], tmp + start[c], end[c] - start[c] + 1);
        int64_t min_chunk = n;
        for (int c = 0; c < sections; c++)
            if (end[c] - start[c] + 1 < min_chunk)
                min_chunk = end[c] - start[c] + 1;
        int64_t ov = min_chunk / overlap_div;
        if (ov < min_overlap)
            ov = min_overlap;
        if (ov <= 0)
            ov = 1;
        /* seams in order (the reference runs them concurrently; the windows are
         * disjoint whenever min_chunk/8 >= 50, i.e. for n >= 9600) */
        for (int s = 0; s < sections - 1; s++) {
            int64_t ls = end[s] - ov + 1;
            if (ls < start[s])
                ls = start[s];
            int64_t re = start[s + 1] + ov - 1;
            if (re > end[s + 1])
                re = end[s + 1];
            if (ls <= re)
                stable_sort_items(items + ls, tmp + ls, re - ls + 1);
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        perm[i] = items[i].idx;
    free(items);
    free(tmp);
}

/* ------------------------------------------------------------------- step */

/* nsphere.c:5796-5838: r_crit switch and per-particle integration at rank i. */
static inline uint8_t integrate_one(int64_t i, int64_t npts, double *r_io, double *v_io,
                                    double ell, const orc_params *p)
{
    double r = *r_io, v = *v_io;
    double r_crit = 0.0;
    if (fabs(ell) > 1.0e-30) {
        double M_enc;
        if (i == 0)
            M_enc = p->menc ? 0.1 * p->menc[npts + 1] : 0.1 * (1.0 / (double)npts) * p->halo_mass;
        else
            M_enc = p->menc ? p->menc[i] : ((double)i / (double)npts) * p->halo_mass;
        double grav_part = (p->vel_conv_sq * p->G) * M_enc;
        r_crit = (ell * ell) * p->alpha / grav_part;
    }
    double r_new, v_new;
    int rec;
    if (r > 1.0e-30 && r < r_crit)
        rec = orc_adaptive_full_lc((int)i, (int)npts, r, v, ell, p->dt, p, &r_new, &v_new, NULL);
    else
        rec = orc_adaptive_full_leap((int)i, (int)npts, r, v, ell, p->dt, p, &r_new, &v_new);
    *r_io = r_new;
    *v_io = v_new;
    return (uint8_t)rec;
}

void orc_integrate(double *r, double *v, const double *ell, int64_t npts,
                   const orc_params *p, uint8_t *level)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < npts; i++) {
        uint8_t rec = integrate_one(i, npts, &r[i], &v[i], ell[i], p);
        if (level)
            level[i] = rec;
    }
}

static void permute_d(double *a, const uint32_t *perm, int64_t n, double *tmp)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        tmp[i] = a[perm[i]];
    memcpy(a, tmp, (size_t)n * sizeof(double));
}

void orc_step(double *r, double *v, double *ell, uint32_t *id, int64_t npts,
              const orc_params *p, int chunked_sort, int tiebreak_id,
              int32_t *inverse_map, uint8_t *level)
{
    uint32_t *perm = (uint32_t *)malloc((size_t)(npts > 0 ? npts : 1) * sizeof(uint32_t));
    double *tmp = (double *)malloc((size_t)(npts > 0 ? npts : 1) * sizeof(double));
    /* sort, nsphere.c:5784-5787 (all columns move together) */
    if (chunked_sort)
        orc_sort_perm_chunked(r, npts, perm);
    else
        orc_sort_perm(r, id, npts, tiebreak_id, perm);
    permute_d(r, perm, npts, tmp);
    permute_d(v, perm, npts, tmp);
    permute_d(ell, perm, npts, tmp);
    {
        uint32_t *itmp = (uint32_t *)tmp;
        uint32_t *ibuf = (uint32_t *)malloc((size_t)(npts > 0 ? npts : 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < npts; i++)
            ibuf[i] = id[perm[i]];
        memcpy(id, ibuf, (size_t)npts * sizeof(uint32_t));
        free(ibuf);
        (void)itmp;
    }
    /* inverse map, nsphere.c:5789-5794 */
    if (inverse_map) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < npts; i++)
            inverse_map[id[i]] = (int32_t)i;
    }
    orc_integrate(r, v, ell, npts, p, level);
    free(perm);
    free(tmp);
}

/* Neumaier-compensated long double sum (error ~2^-64 relative, far below the
 * one final rounding to double); needs the strict build (no reassociation). */
void orc_mass_prefix(const double *mass, int64_t npts, double *out)
{
    long double s = 0.0L, c = 0.0L;
    for (int64_t i = 0; i < npts; i++) {
        out[i] = (double)(s + c);
        const long double x = (long double)mass[i];
        const long double t = s + x;
        if (fabsl(s) >= fabsl(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    out[npts] = (double)(s + c);
    out[npts + 1] = npts > 0 ? mass[0] : 0.0;
}

void orc_energy_masses(const double *r, const double *v, const double *ell, const double *mass,
                       int64_t npts, const orc_params *p, double *K, double *W)
{
    if (!mass) {
        orc_energy(r, v, ell, npts, p, K, W);
        return;
    }
    const long double gpre = (long double)p->vel_conv_sq * (long double)p->G;
    long double ks = 0.0L, ws = 0.0L, menc = 0.0L;
    for (int64_t i = 0; i < npts; i++) {
        long double rr = r[i], vv = v[i], ll = ell[i], mi = mass[i];
        ks += mi * 0.5L * (vv * vv + (ll * ll) / (rr * rr));
        if (!p->identity_gravity)
            ws -= gpre * menc * mi / rr;
        menc += mi;
    }
    *K = (double)ks;
    *W = (double)ws;
}

void orc_energy(const double *r, const double *v, const double *ell, int64_t npts,
                const orc_params *p, double *K, double *W)
{
    const double m = p->halo_mass / (double)npts;
    const double gpre = p->vel_conv_sq * p->G;
    long double ks = 0.0L, ws = 0.0L;
    for (int64_t i = 0; i < npts; i++) {
        long double rr = r[i], vv = v[i], ll = ell[i];
        ks += (long double)m * 0.5L * (vv * vv + (ll * ll) / (rr * rr));
        long double menc = ((long double)i / (long double)npts) * (long double)p->halo_mass;
        if (!p->identity_gravity)
            ws -= (long double)gpre * menc * (long double)m / rr;
    }
    *K = (double)ks;
    *W = (double)ws;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}
