// Host-visible launch interface of the tracer advection kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace fesomtra {

struct MeshDev;

// Arguments of one tracer step (device pointers, device numbering).
struct StepArgs {
    int hor, ver;        // Hor / Ver codes
    bool fct;            // limiter "FCT" (else "NONE")
    double ph, pv, dt;   // tra_adv_ph, tra_adv_pv, time step
    double *T;           // values (updated in place unless acc_h / acc_v are set)
    const double *TAB;   // valuesAB (== T for the caller step)
    double *flo, *adfh;  // (E, L) work
    double *LO, *Rp, *Rm;  // (N, L) work
    double *adfv;        // (N, NL) work
    double *acc_h, *acc_v;  // do_oce_adv_tra accumulators or null
    double *dbg_h, *dbg_v;  // limited fluxes for inspection or null
    hipEvent_t *ev;         // optional: 5 events recorded around the 4 stages, or null
};

hipError_t launch_fct_step(const MeshDev &m, const StepArgs &a, hipStream_t s);
hipError_t launch_gather_rows(const double *in, double *out, const int *map, int rows, int ld,
                              hipStream_t s);
hipError_t launch_stats(const MeshDev &m, const double *T, int all_levels, double4 *part,
                        int nblocks, hipStream_t s);

}  // namespace fesomtra
