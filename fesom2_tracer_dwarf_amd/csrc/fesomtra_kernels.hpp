// Host-visible launch interface of the tracer advection kernels.
#pragma once

#include "fesomtra_types.hpp"

namespace fesomtra {

struct SchemeArgs {
    int hor, ver;       // Hor / Ver codes
    bool fct;           // limiter "FCT" (else "NONE")
    double ph, pv, dt;  // tra_adv_ph, tra_adv_pv, time step
};

// Levels per thread of each kernel (the column of a node / edge is split into chunks of lc
// layers, one thread per chunk).
struct LaunchCfg {
    int lc_edge = 16, lc_lo = 16, lc_bounds = 16, lc_final = 16;
};

// One step of kt (1 or 2) tracers sharing one scheme.  ev, when not null, holds 5 events recorded
// around the 4 stages (edge fluxes, low-order / vertical, bounds, final).
template <typename S, typename M, typename C>
hipError_t launch_step(const MeshCol<M> &m, const StepPtrs<S, C> &p, int kt, const SchemeArgs &sc,
                       const LaunchCfg &cfg, hipStream_t s, hipEvent_t *ev);

// dir 0: reference layout (nl-1, nod2D) float64 -> device layout S; dir 1: back.
template <typename S>
hipError_t launch_relayout(const double *ref_in, S *dev_out, const S *dev_in, double *ref_out,
                           const int *perm, int N, int L, size_t Np, int dir, hipStream_t s);

template <typename S, typename M>
hipError_t launch_stats(const MeshCol<M> &m, const S *T, int all_levels, double4 *part, int nblocks,
                        hipStream_t s);

}  // namespace fesomtra
