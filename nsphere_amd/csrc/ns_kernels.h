// ns_kernels.h -- launch interface of the integrate / gather / energy kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ns {

// Everything the fused step kernel needs, passed by value.
struct StepArgs {
    const double *r0, *v0, *l0;  // state in the previous rank order
    const uint32_t *id0;
    const uint32_t *perm;        // new rank -> previous position (GATHER form)
    double *r1, *v1, *l1;        // state in the new rank order
    uint32_t *id1;
    uint64_t *keys;              // radius keys of the new r (may be null)
    uint8_t *lvl;                // level records (may be null)
    unsigned long long *budget_hits;
    int64_t n;
    double dt, G, halo_mass, vel_conv_sq, radius_tol, velocity_tol, alpha;
    int64_t max_subdiv;
    int64_t budget;              // force evaluations allowed per particle per step
    int32_t out_type, N_taumin, identity_gravity;
    // binned integrator: queues of ranks handed between its kernels
    uint32_t *q_lf;              // leapfrog particles past cap_N (wave per particle)
    uint32_t *q_lc;              // Levi-Civita particles (lane per particle)
    uint32_t *q_lcw;             // Levi-Civita particles past cap_lc in some tau step (wave)
    uint32_t *q_count;           // lengths of q_lf, q_lc, q_lcw
    int64_t cap_N;               // deepest leapfrog N tried lane-per-particle
    int64_t cap_lc;              // deepest Levi-Civita N tried lane-per-particle
};

// Integrator forms (NS_OPT_INTEGRATOR).
enum IntegratorForm : int {
    FORM_PLAIN = 0,        // one lane per particle, sequential levels, strict IEEE
    FORM_BINNED = 1,       // bulk lanes + deep queue with wave-parallel levels, strict IEEE
    FORM_BINNED_FAST = 2   // as BINNED with the fast arithmetic form
};

hipError_t launch_integrate(const StepArgs &a, bool gather, int form, hipStream_t stream);
hipError_t launch_gather_state(const double *r0, const double *v0, const double *l0, const uint32_t *id0,
                               const uint32_t *perm, double *r1, double *v1, double *l1, uint32_t *id1,
                               int64_t n, hipStream_t stream);
hipError_t launch_inverse_map(const uint32_t *id, int32_t *inv, int64_t n, hipStream_t stream);
int energy_blocks(int64_t n);
hipError_t launch_energy(const double *r, const double *v, const double *ell, const uint32_t *perm,
                         int64_t n, double m, double gpre, double halo_mass, int idg, double *partial,
                         double *out, hipStream_t stream);

} // namespace ns
