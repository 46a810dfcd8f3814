// Device data layout and precision policy of the tracer advection kernels.
//
// Layout in HBM: level-major ("one 2-D field per level"), entities renumbered along a Hilbert
// curve and padded to a multiple of 64 per level:
//   node-layer      a[k * Np + n]      k = nz-1 in [0, nl-1)
//   node-interface  a[j * Np + n]      j = interface-1 in [0, nl)
//   element-layer   a[k * Tp + el],    velocities uv[(k * Tp + el) * 2 + c]
//   edge-layer      a[k * Ep + e]
// Tracers and per-tracer work arrays are stored in blocks of KB tracers interleaved innermost:
//   tracer block    t[(k * Np + n) * KB + lane]     (edge work arrays with Ep in place of Np)
// One thread owns one (column, lane) pair and walks the column's wet levels; the KB lanes of a
// column are adjacent threads, so mesh values are fetched once per column for all KB tracers and
// every tracer access is a contiguous run of KB values over consecutive columns.  Narrow blocks
// (KB <= 4) work on tiles of 256 / KB consecutive columns or edges whose neighbour values are
// staged in LDS level by level (MeshCol::ht_*, MeshCol::et_*).
//
// Precision policy: S = tracer storage, M = mesh / velocity values (double), C = high-order
// reconstruction arithmetic and the stored intermediates.  The edge transport q, the low-order
// fluxes and solution, the vertical fluxes, the nodal flux-divergence sums, the limiter ratios and
// the tracer update always run in double.  With S = C = double the kernels keep the reference's
// operation order and are bit-identical to the serial FP64 reference; with C = float the
// antidiffusive edge flux is evaluated in an algebraically equivalent difference form (exactly
// zero for a uniform tracer) so that single-precision rounding does not break the discrete
// cancellation of a divergence-free transport.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace fesomtra {

constexpr double kREarth = 6367500.0;  // lib/oce_modules.F90:24
constexpr int kXcds = 8;               // MI355X: 8 accelerator complex dies, one L2 each
constexpr int kMaxKB = 64;             // widest tracer block (one wavefront of lanes)

enum Hor { HOR_UPW1 = 0, HOR_MUSCL = 1, HOR_MFCT = 2 };
enum Ver { VER_UPW1 = 0, VER_QR4C = 1, VER_CDIFF = 2, VER_PPM = 3 };
enum Prec { PREC_F64 = 0, PREC_F32 = 1, PREC_F32_F16 = 2, PREC_F32_BF16 = 3 };

typedef __half f16_t;
typedef __hip_bfloat16 bf16_t;

// Read-only mesh data: loads through the constant address space, so that loads at a
// wavefront-uniform index become scalar loads (one per wavefront into scalar registers); the
// arrays are never written while a kernel that reads them runs.
template <class T>
__device__ __forceinline__ T cld(const T *p, size_t i) {
    return ((const __attribute__((address_space(4))) T *)p)[i];
}
__device__ __forceinline__ int4 cld(const int4 *p, size_t i) {
    const int *q = reinterpret_cast<const int *>(p) + 4 * i;
    return make_int4(cld(q, 0), cld(q, 1), cld(q, 2), cld(q, 3));
}
__device__ __forceinline__ int2 cld(const int2 *p, size_t i) {
    const int *q = reinterpret_cast<const int *>(p) + 2 * i;
    return make_int2(cld(q, 0), cld(q, 1));
}

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

// Stochastic rounding of the tracer update into single / half storage: rounds v up with
// probability (v - down) / (up - down), with `u` uniform in [0, 1), so the stored value is exact
// in expectation and slowly varying fields neither stagnate nor drift.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ double unit_hash(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t h = mix32(a * 0x9e3779b1u ^ mix32(b * 0x7feb352du ^ mix32(c + 0x68e31da4u)));
    return (double)(h >> 8) * (1.0 / 16777216.0);
}
__device__ __forceinline__ float bf16_bits_to_float(uint32_t b) { return __uint_as_float(b << 16); }
template <typename T>
__device__ __forceinline__ void st_stochastic(T *__restrict__ p, size_t i, double v, double u) {
    if constexpr (std::is_same<T, double>::value) {
        p[i] = v;
    } else if constexpr (std::is_same<T, float>::value) {
        const float lo = __double2float_rd(v), hi = __double2float_ru(v);
        p[i] = (lo == hi || u >= (v - (double)lo) / ((double)hi - (double)lo)) ? lo : hi;
    } else if constexpr (std::is_same<T, f16_t>::value) {
        const __half lo = __float2half_rd(__double2float_rd(v)), hi = __float2half_ru(__double2float_ru(v));
        const double dl = (double)__half2float(lo), dh = (double)__half2float(hi);
        p[i] = (dl == dh || u >= (v - dl) / (dh - dl)) ? lo : hi;
    } else {
        const float f_lo = __double2float_rd(v), f_hi = __double2float_ru(v);
        const uint32_t bl = __float_as_uint(f_lo), bh = __float_as_uint(f_hi);
        // bf16 neighbours: truncate the float bit patterns toward -inf / +inf
        const uint32_t dn = (bl & 0xffffu) == 0u ? bl >> 16 : (f_lo >= 0.f ? bl >> 16 : (bl >> 16) + 1u);
        const uint32_t up = (bh & 0xffffu) == 0u ? bh >> 16 : (f_hi >= 0.f ? (bh >> 16) + 1u : bh >> 16);
        const double dl = (double)bf16_bits_to_float(dn), dh = (double)bf16_bits_to_float(up);
        const uint32_t pick = (dl == dh || u >= (v - dl) / (dh - dl)) ? dn : up;
        p[i] = __float2bfloat16(bf16_bits_to_float(pick));
    }
}

// a / b for a divisor used many times, given y = 1 / b (correctly rounded): q0 = a y is within
// one unit in the last place of a / b, the residual a - b q0 is exact with a fused multiply-add,
// and q0 + (a - b q0) y is the correctly rounded quotient (P. W. Markstein, IBM J. Res. Dev. 34,
// 1990).  A zero numerator returns a y (the signed zero of a / b); quotients near the underflow
// or overflow thresholds, and non-finite ones, take the division itself.  The result is
// bit-identical to a / b.
__device__ __forceinline__ double div_reuse(double a, double b, double y) {
    const double q0 = a * y;
    if (a == 0.0) return q0;
    const double aq = fabs(q0);
    if (aq > 0x1p-960 && aq < 0x1p960) return fma(fma(-q0, b, a), y, q0);
    return a / b;
}

template <typename C>
__device__ __forceinline__ C cmin(C a, C b) { return a < b ? a : b; }
template <typename C>
__device__ __forceinline__ C cmax(C a, C b) { return a > b ? a : b; }
template <typename C>
__device__ __forceinline__ C csign1(C x) { return signbit(x) ? C(-1) : C(1); }

// Mesh arrays on the device (device numbering; levels are 1-based in the formulas).
template <typename M>
struct MeshCol {
    int N, T, E, nl, L, Np, Tp, Ep, max_adj;
    // Halo tiles of the node kernels for tracer blocks of width ht_kb (<= 4; ht_n = 0: none):
    // tile t owns the 256 / ht_kb consecutive nodes from t * 256 / ht_kb and lists its one-ring
    // halo; ht_loc gives, per node -> edge entry, the far node's slot in its tile (tile column,
    // or 256 / ht_kb + halo position).
    int ht_n, ht_kb;
    const int *ht_loc, *ht_hptr, *ht_halo, *ht_lmax;  // (entries), (ht_n + 1), (halo), (ht_n) wet layers
    // The edges of each tile's nodes, sorted (ht_eptr, ht_edges; at most 1024 per tile at width 1,
    // 512 at widths 2 and 4, else no tiles are built), and per CSR entry its edge's position in
    // that list * 2 + whether the node is the edge's second end (ht_eloc): the low-order kernel
    // stages the tile's transports level by level.
    const int *ht_eptr, *ht_edges, *ht_eloc;
    // Edge tiles of the edge-flux kernel (same widths): tile t owns the 256 / et_kb consecutive
    // edges from t * 256 / et_kb and lists the nodes they read (ends and up/down-wind triangle
    // corners); et_loc holds per edge the slots of n1, n2, u0, u1, u2, d0, d1, d2 (16 bits each).
    int et_n, et_kb;
    const int *et_loc, *et_nptr, *et_nodes, *et_lmax;  // (E, 4), (et_n + 1), (nodes), (et_n) layers
    // Tiles of the fused limiter kernel for the same widths (rt_n = 0: none, or a tile's rings
    // did not fit): tile t owns the 128 / ht_kb consecutive nodes from t * 128 / ht_kb; rt_halo
    // lists its first ring (rt_nr1[t] nodes, at most 128 / ht_kb) and then its second ring (at
    // most 256 / ht_kb); per node thread i of the tile (tile node, then first-ring node) rt_eoff[t *
    // 256 / ht_kb + i] is the offset in rt_loc of its CSR entries' far nodes as staging slots
    // (tile node k: k, first-ring node k: 128 / ht_kb + k, second-ring node k: 256 / ht_kb + k).
    // rt_lmax: the tile nodes' deepest layer; rt_smax: that of the tile and its first ring.
    int rt_n;
    const int *rt_hptr, *rt_halo, *rt_nr1, *rt_eoff, *rt_loc, *rt_lmax, *rt_smax;
    // nodes
    const int *nlev_n, *nlev_min, *nbl;   // nlevels_nod2D, nlevels_nod2D_min, nboundary_lay
    const int *ne_ptr;                    // CSR node -> edges, reference edge order
    const int *ne_e2;                     // edge * 2 + (node is edges(2))
    const int *ne_top;                    // deepest level of the edge's scatter range
    const int *ne_other;                  // other end node
    const int *nie, *nie_num;             // nod_in_elem2D (N, max_adj), reference list order
    const M *hnode, *hnode_new, *Z;       // [L][Np]
    const M *area, *areasvol, *zbar;      // [nl][Np]
    const M *w, *we, *wi;                 // [nl][Np] w, its explicit and implicit parts
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
    double *q;                            // [L][Ep] edge transport of the current velocity
    // Ensemble members (several velocities on this mesh; nmem <= 1: one): the velocity arrays hold
    // nmem members, uv member-major ([nmem][L][Tp][2]), w / we / wi and q member-interleaved
    // ([nl][Np][nmem], [L][Ep][nmem]); each lane of a tracer block reads its own member's values.
    int nmem;
    // Node-mean gradient entries: a MUSCL / MFCT edge at a level outside its up/down triangles'
    // shared wet range takes the area-weighted mean gradients of both end nodes.  Entry
    // mg_base[n] + nz is node n at level nz, for the levels any of its edges asks for; mg_node /
    // mg_lev list the mg_n entries.  Evaluated once per pass by k_mgrad (StepPtrs::mg).
    int mg_n;
    const int *mg_node, *mg_lev, *mg_base;
    const double *dtn;                    // (N) time-step factor per node, or null (all 1)
    // Ice-shelf cavities (layers above a top are not wet); all null without them (tops are 1)
    const int *ulev_n;                    // (N) ulevels_nod2D, the node's top layer
    const int *ulev_max;                  // (N) ulevels_nod2D_max, deepest element top around the node
    const int *ulev_e;                    // (T) ulevels, the element's top layer
    const int *ne_utop;                   // CSR entries: top level of the edge's scatter range
    // The limiter's cluster bound reads, for an element whose top lies below the level, the work
    // array it shares with the edge gradients (lib/oce_adv_tra_fct.F90:77,164-184): at element
    // number el that holds the gradient fill of the edge with the same reference number.  Slots
    // are those elements: the device edge with the element's reference number (cq_edge), the
    // element's top (cq_top) and the offset of its rows, one per level above the top, in the
    // per-pass values StepPtrs::cqv (cq_off).  Per node, the slots of the elements around it
    // whose top lies below the node's (cq_ptr, cq_ent).
    int cq_n;
    const int *cq_edge, *cq_top, *cq_off, *cq_ptr, *cq_ent;
};

// Top layer of node n's column (1 without cavities; GEN: kernels for cavity meshes).
template <bool GEN = true, typename M>
__device__ __forceinline__ int node_top(const MeshCol<M> &m, int n) {
    if constexpr (GEN) return m.ulev_n ? cld(m.ulev_n, n) : 1;
    else return 1;
}
// Top layer of element el.
template <typename M>
__device__ __forceinline__ int elem_top(const MeshCol<M> &m, int el) {
    return m.ulev_e ? cld(m.ulev_e, el) : 1;
}

// Time step of node n's cells: dt, or dt * dtn[n] where the mesh carries a factor per node (a
// union of independent meshes, each advanced with its own time step).  GEN: kernels compiled for
// meshes with such factors or with cavities (the others keep dt a scalar kernel argument).
template <bool GEN, typename M>
__device__ __forceinline__ double node_dt(const MeshCol<M> &m, int n, double dt) {
    if constexpr (GEN) return m.dtn ? dt * (double)cld(m.dtn, n) : dt;
    else return dt;
}

// Member of a lane and the member count, the index and stride into the member-interleaved
// velocity arrays: a[i * n + m] for entry i of the single-velocity layout.
struct Member {
    int m = 0, n = 1;
    __device__ __forceinline__ size_t at(size_t i) const { return i * (size_t)n + (size_t)m; }
};

// Pointers of one step over one tracer block (arrays interleaved by the block width KB).
template <typename S, typename C>
struct StepPtrs {
    S *T;                  // values (updated in place unless accumulating, or unless Tout is set)
    S *Tout;               // fused limiter (k_limit): the updated values go here, T is left as read
                           // (neighbouring tiles read T of their halos while the update runs), or null
    const S *TAB;          // valuesAB (may equal T)
    C *adfh;               // [L][Ep][KB] antidiffusive (or high-order) edge flux
    C *LO;                 // [L][Np][KB] low-order solution (single precision: its increment LO - T)
    C *LOe;                // single precision with compensated storage: the increment's remainder, or null
    C *Rp, *Rm;            // [L][Np][KB] limiter ratios
    C *adfv;               // [nl][Np][KB] antidiffusive (or high-order) vertical flux
    double *acc_h, *acc_v; // do_oce_adv_tra accumulators [L][Np] of the single selected lane, or null
    int acc_init;          // the accumulators are logically zero (zeroed since the last call): the
                           // update starts its sums from 0 instead of reading them
    C *dbg_h, *dbg_v;      // limited fluxes [L][Ep], [nl][Np] of the single selected lane, or null
    double2 *tri;          // [L][Np][KB] (c', t') of the implicit vertical solve, or null
    S *ab_next, *ab_old;   // AB2: valuesAB of the next step and valuesold, written by the update, or null
    float *resid;          // compensated storage: [L][Np][KB] FP32 remainder of the stored values, or null
    double ab_c0, ab_c1;   // 0.5 + epsilon, 1.5 + epsilon
    C *cqv;                // cavity meshes: [rows][2][KB] the limiter's element-bound values above an
                           // element's top (MeshCol::cq_*), written by k_cavity_aux
    unsigned long long mask;  // lanes advanced by this launch (tracers sharing its scheme)
    int tid0;                 // tracer number of lane 0 (seeds the stochastic rounding)
    uint32_t seed;            // nonzero: stochastic rounding of the update; 0: round to nearest
    const uint32_t *seed_dev; // the step counter that seeds the stochastic rounding, in device
                              // memory (advanced by a kernel per step, so captured graphs replay
                              // with new seeds); null: `seed` itself is the counter
    const int *lane_mem;      // ensembles: the member of each lane of this block (KB entries), or null
    C *mg;                    // [mg_n][2][KB] node-mean gradients (x, y) of valuesAB, or null
                              // (then the edge kernel evaluates them where it needs them)
};

// Velocity member of `lane` (GEN: kernels for ensembles, per-node time steps or cavities; the
// others read the single-velocity layout).
template <bool GEN, typename M, typename S, typename C>
__device__ __forceinline__ Member lane_member(const MeshCol<M> &m, const StepPtrs<S, C> &p, int lane) {
    if constexpr (GEN) return p.lane_mem ? Member{cld(p.lane_mem, (size_t)lane), m.nmem} : Member{0, 1};
    else return Member{0, 1};
}

}  // namespace fesomtra
