// Host-visible launch interface of the tracer advection kernels.
#pragma once

#include "fesomtra_types.hpp"

namespace fesomtra {

struct SchemeArgs {
    int hor, ver;         // Hor / Ver codes
    bool fct;             // limiter "FCT" (else "NONE")
    double ph, pv, dt;    // tra_adv_ph, tra_adv_pv, time step
    bool wsplit = false;  // dynamics%use_wsplit (FCT only): implicit low-order part with w_i
    bool cavity_fill = false;  // cavity meshes: the edge gradients were filled for this tracer
};

// Levels per thread of each kernel (the column of a node / edge is split into chunks of lc
// layers, one thread per chunk).
struct LaunchCfg {
    int lc_edge = 16, lc_lo = 16, lc_bounds = 16, lc_final = 16;
    // With the limiter the low-order kernels run on `aux` beside the edge-flux kernels: `fork` is
    // recorded on the main stream before them and `join` on aux after them (null: one stream).
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// One step of the lanes p.mask of a tracer block of width kb (1, 2, 4, ..., 64); the selected
// lanes share the scheme sc.  ev, when not null, holds 5 events recorded around the 4 stages
// (edge fluxes, low-order / vertical, bounds, final).
// Edge transport q = m.q of the current velocity (once per step, before the tracer passes).
template <typename M>
hipError_t launch_qflux(const MeshCol<M> &m, hipStream_t s);

// Step counter of the stochastic rounding in device memory: advance = 0 sets it to v, 1 adds one
// (skipping 0).  Built with every precision family's object only once (FP64 family).
hipError_t launch_seed(uint32_t *c, uint32_t v, int advance, hipStream_t s);

template <typename S, typename M, typename C>
hipError_t launch_step(const MeshCol<M> &m, const StepPtrs<S, C> &p, int kb, const SchemeArgs &sc,
                       const LaunchCfg &cfg, hipStream_t s, hipEvent_t *ev);

// dir 0: reference layout (nl-1, nod2D) float64 -> lane of a device tracer block; dir 1: back.
// res (compensated storage, FP32, or null): dir 0 also stores the rounding remainder there, dir 1
// adds it.
template <typename S>
hipError_t launch_relayout(const double *ref_in, S *dev_out, const S *dev_in, double *ref_out,
                           const int *perm, int N, int L, size_t Np, int kb, int lane, int dir, hipStream_t s,
                           float *res = nullptr);

// valuesAB = -(0.5+eps)*valuesold + (1.5+eps)*values and valuesold := values over n entries of a
// tracer block (FESOM2's AB2 interpolation).
template <typename S>
hipError_t launch_ab2(const S *T, S *AB, S *old, size_t n, double eps, hipStream_t s);

template <typename S, typename M>
hipError_t launch_stats(const MeshCol<M> &m, const S *T, int kb, int lane, int all_levels, double4 *part,
                        int nblocks, hipStream_t s, const float *res = nullptr);

}  // namespace fesomtra
