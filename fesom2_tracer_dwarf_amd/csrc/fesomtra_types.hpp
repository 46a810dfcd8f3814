// Device data layout and precision policy of the tracer advection kernels.
//
// Layout in HBM: level-major ("one 2-D field per level"), entities renumbered along a Hilbert
// curve and padded to a multiple of 64 per level:
//   node-layer      a[k * Np + n]      k = nz-1 in [0, nl-1)
//   node-interface  a[j * Np + n]      j = interface-1 in [0, nl)
//   element-layer   a[k * Tp + el],    velocities uv[(k * Tp + el) * 2 + c]
//   edge-layer      a[k * Ep + e]
// One thread owns one column (node, element or edge) and walks its wet levels, so per-entity
// data is loaded once per column and every per-level load is coalesced over consecutive entities.
//
// Precision policy: S = tracer storage, M = mesh / velocity values, C = flux and reconstruction
// arithmetic and the stored intermediates; the nodal flux-divergence sums, the low-order
// solution, the limiter ratios and the tracer update always run in double.  With
// S = M = C = double the kernels keep the reference's operation order and are bit-identical to
// the serial FP64 reference.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fesomtra {

constexpr double kREarth = 6367500.0;  // lib/oce_modules.F90:24
constexpr int kXcds = 8;               // MI355X: 8 accelerator complex dies, one L2 each
constexpr int kMaxKT = 2;              // tracers per thread in the column kernels
constexpr int kDegCache = 8;           // node->edge entries held in registers

enum Hor { HOR_UPW1 = 0, HOR_MUSCL = 1, HOR_MFCT = 2 };
enum Ver { VER_UPW1 = 0, VER_QR4C = 1, VER_CDIFF = 2, VER_PPM = 3 };
enum Prec { PREC_F64 = 0, PREC_F32 = 1, PREC_F32_F16 = 2, PREC_F32_BF16 = 3 };

typedef __half f16_t;
typedef __hip_bfloat16 bf16_t;

// Loads converting to the compute type C, stores rounding from double.
template <typename C, typename T>
__device__ __forceinline__ C ldc(const T *__restrict__ p, size_t i) {
    if constexpr (std::is_same<T, f16_t>::value) return (C)__half2float(p[i]);
    else if constexpr (std::is_same<T, bf16_t>::value) return (C)__bfloat162float(p[i]);
    else return (C)p[i];
}
template <typename T>
__device__ __forceinline__ void std_(T *__restrict__ p, size_t i, double v) {
    if constexpr (std::is_same<T, f16_t>::value) p[i] = __float2half_rn((float)v);
    else if constexpr (std::is_same<T, bf16_t>::value) p[i] = __float2bfloat16((float)v);
    else p[i] = (T)v;
}

template <typename C>
__device__ __forceinline__ C cmin(C a, C b) { return a < b ? a : b; }
template <typename C>
__device__ __forceinline__ C cmax(C a, C b) { return a > b ? a : b; }
template <typename C>
__device__ __forceinline__ C csign1(C x) { return signbit(x) ? C(-1) : C(1); }

// Mesh arrays on the device (device numbering; levels are 1-based in the formulas).
// ulevels == 1 everywhere (cavities are rejected at setup).
template <typename M>
struct MeshCol {
    int N, T, E, nl, L, Np, Tp, Ep, max_adj;
    // nodes
    const int *nlev_n, *nlev_min, *nbl;   // nlevels_nod2D, nlevels_nod2D_min, nboundary_lay
    const int *ne_ptr;                    // CSR node -> edges, reference edge order
    const int *ne_e2;                     // edge * 2 + (node is edges(2))
    const int *ne_top;                    // deepest level of the edge's scatter range
    const int *ne_other;                  // other end node
    const int *nie, *nie_num;             // nod_in_elem2D (N, max_adj), reference list order
    const M *hnode, *hnode_new, *Z;       // [L][Np]
    const M *area, *areasvol, *zbar;      // [nl][Np]
    const M *w, *we;                      // [nl][Np]
    // elements
    const int *enodes, *nlev_e;           // (T, 3), (T)
    const M *elem_area, *elem_cos, *gsca; // (T), (T), (T, 6)
    const M *helem;                       // [L][Tp]
    const M *uv;                          // [L][Tp][2]
    // edges
    const int4 *edge_ne;                  // n1, n2, el1, el2 (-1 none)
    const int2 *updn;                     // up / down triangle (-1 none)
    const M *ecd;                         // (E, 4) edge_cross_dxdy
    const M *edxdy;                       // (E, 2) edge_dxdy
};

// Per-slot (tracer within a thread) pointers of one step.
template <typename S, typename C>
struct StepPtrs {
    S *T[kMaxKT];          // values (updated in place unless accumulating)
    const S *TAB[kMaxKT];  // valuesAB (may equal T)
    C *flo[kMaxKT];        // (E) low-order edge flux
    C *adfh[kMaxKT];       // (E) antidiffusive (or high-order) edge flux
    C *LO[kMaxKT];         // (N) low-order solution
    C *Rp[kMaxKT], *Rm[kMaxKT];  // (N) limiter ratios
    C *adfv[kMaxKT];       // (N interfaces) antidiffusive (or high-order) vertical flux
    double *acc_h[kMaxKT], *acc_v[kMaxKT];  // do_oce_adv_tra accumulators (or null)
    C *dbg_h[kMaxKT], *dbg_v[kMaxKT];       // limited fluxes for inspection (or null)
};

}  // namespace fesomtra
