// gfx950 kernels of the tracer advection step and their launchers.
//
// One step of a block of KB interleaved tracers (the lanes of `mask` share one scheme) with the
// FCT limiter is four column kernels (do_oce_adv_tra, lib/oce_adv_tra_driver.F90:23-294):
//   k_edge    one thread per edge and lane: high-order minus low-order (upwind) flux for every
//             level of the edge; MUSCL/MFCT gradients reconstructed on the fly from the
//             up/down-wind triangles (no edge_up_dn_grad array)
//   k_lo      one thread per node and lane: deterministic gather of the edge fluxes, vertical upwind
//             flux, low-order solution fct_LO, vertical antidiffusive flux (QR4C/PPM/...)
//   k_bounds  one thread per node and lane: cluster bounds over the node's 1-ring, vertical widening,
//             P+/P- gathers and the limiter ratios R+/R- (lib/oce_adv_tra_fct.F90:96-374)
//   k_final   one thread per node and lane: limiting with R at both ends, flux divergence, tendencies
//             and the tracer update in place (lib/oce_adv_tra_fct.F90:389-471,
//             lib/oce_adv_tra_driver.F90:298-450, src/fesom_analytic.F90:419-437)
// Without FCT: k_edge, k_vert and k_final.  No atomics: each node adds its incident edges in
// the reference's edge order, so sums are deterministic and equal the serial reference.
// For tracer blocks of width <= 4 the kernels run on tiles (EdgeTile for k_edge, HaloTile for the
// three FCT node kernels): neighbour values are staged in LDS per level instead of being
// gathered once per edge.  A node's sum over its edges issues the loads of all its edges before
// consuming them (for_edges_ld), so a level waits for one memory latency rather than one per edge.
#include "fesomtra_column.hpp"
#include "fesomtra_kernels.hpp"

namespace fesomtra {

// Occupancy target (waves per SIMD) handed to the register allocator: columns of one lane per
// thread are latency-bound gathers and gain from more resident waves.
#ifndef FESOMTRA_MIN_WAVES
#define FESOMTRA_MIN_WAVES 4
#endif
template <int KB>
constexpr int kMinWaves = KB == 64 ? 1 : FESOMTRA_MIN_WAVES;
// Halo-tiled node kernels hold the staging registers as well; their workgroups also share LDS.
#ifndef FESOMTRA_MIN_WAVES_TILED
#define FESOMTRA_MIN_WAVES_TILED 4
#endif
template <bool TILED, int KB>
constexpr int kWaves = TILED ? FESOMTRA_MIN_WAVES_TILED : kMinWaves<KB>;
// The tiled edge kernel's own target (it holds a group of levels in flight).
#ifndef FESOMTRA_MIN_WAVES_EDGE
#define FESOMTRA_MIN_WAVES_EDGE FESOMTRA_MIN_WAVES_TILED
#endif
template <bool TILED, int KB>
constexpr int kEdgeWaves = TILED ? FESOMTRA_MIN_WAVES_EDGE : kMinWaves<KB>;

// ---------------------------------------------------------------------------------------------
// Work decomposition: one thread owns `lc` consecutive levels of one column for one lane of the
// tracer block; blockIdx.x selects columns (XCD-aware, see column_index) and blockIdx.y the level
// chunk.  Vertical stencils that
// cross a chunk boundary recompute the few values they need from the neighbouring chunk.

// 1-based layer range [nz0, nz1] of this thread's chunk in a column of L wet layers.
__device__ __forceinline__ bool chunk_layers(int L, int lc, int &nz0, int &nz1) {
    nz0 = (int)blockIdx.y * lc + 1;
    if (nz0 > L) return false;
    nz1 = min(nz0 + lc - 1, L);
    return true;
}

// ---------------------------------------------------------------------------------------------
// Halo tiles (node kernels of tracer blocks of width KB <= 4).  A workgroup of 256 threads owns
// TC = 256 / KB consecutive columns (a compact patch of the space-filling curve) and walks its
// level chunk top-down.  At every level it stages the neighbour quantity of the tile and its
// one-ring halo in LDS (double-buffered; the next level is fetched into registers while the
// current one is computed), so a neighbour value is read from memory once per tile instead of
// once per incident edge.  Thread (col, lane) stages lane `lane` of tile column col (slot 0)
// and of halo node col (slot 1), so a halo holds at most TC nodes; LDS slot of (local node,
// lane) is local * KB + lane.  Arithmetic and summation order are those of the untiled kernels.
constexpr int kHtSlots = 512;

template <int KB, typename M>
struct HaloTile {
    static constexpr int TC = 256 / KB;
    int n0, nz_b, nz_e;  // first column; level range of this chunk (1-based, tile's wet layers)
    int lmax;
    int snode[2];
    bool son[2];
    // false for the whole workgroup when it has nothing to do
    __device__ __forceinline__ bool init(const MeshCol<M> &m, int lc) {
        const int t = xcd_block();
        if (t >= m.ht_n) return false;
        n0 = t * TC;
        lmax = cld(m.ht_lmax, t);
        nz_b = (int)blockIdx.y * lc + 1;
        if (nz_b > lmax) return false;
        nz_e = min(nz_b + lc - 1, lmax);
        const int hb = cld(m.ht_hptr, t), H = cld(m.ht_hptr, t + 1) - hb;
        const int col = (int)threadIdx.x / KB;
        snode[0] = min(n0 + col, m.N - 1);  // columns past the last node: in bounds, never read
        son[0] = true;
        son[1] = col < H;
        snode[1] = son[1] ? cld(m.ht_halo, hb + col) : 0;
        return true;
    }
    // f(i, node) for the staging slots of this thread
    template <typename F>
    __device__ __forceinline__ void each(F &&f) const {
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (son[i]) f(i, snode[i]);
    }
};

// Edge tiles (edge-flux kernel of tracer blocks of width KB <= 4): a workgroup owns TE = 256 / KB
// consecutive edges and, per level, stages the tracer values of every node its edges read (both
// ends, and the corners of the up- and down-wind triangles) in LDS; thread (edge, lane) stages
// lane `lane` of tile nodes edge and TE + edge.  The node-mean gradients of the few levels below
// the shared wet depth of an edge's triangles still read memory.
template <int KB, typename M>
struct EdgeTile {
    static constexpr int TE = 256 / KB;
    int e0, nz_b, nz_e;
    int snode[2];
    bool son[2];
    __device__ __forceinline__ bool init(const MeshCol<M> &m, int lc) {
        const int t = xcd_block();
        if (t >= m.et_n) return false;
        e0 = t * TE;
        const int lmax = cld(m.et_lmax, t);
        nz_b = (int)blockIdx.y * lc + 1;
        if (nz_b > lmax) return false;
        nz_e = min(nz_b + lc - 1, lmax);
        const int nb = cld(m.et_nptr, t), nn = cld(m.et_nptr, t + 1) - nb;
        const int col = (int)threadIdx.x / KB;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            son[i] = col + i * TE < nn;
            snode[i] = son[i] ? cld(m.et_nodes, nb + col + i * TE) : 0;
        }
        return true;
    }
    template <typename F>
    __device__ __forceinline__ void each(F &&f) const {
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (son[i]) f(i, snode[i]);
    }
};
// Reads lane `lane` of staged tile slot `slot`: ld(slot) = buf[slot * KB + lane].
template <int KB, typename C>
struct SlotLd {
    const C *buf;
    int lane;
    __device__ __forceinline__ C operator()(size_t slot) const { return buf[slot * KB + lane]; }
};

#ifndef FESOMTRA_EDGE_PIPE
#define FESOMTRA_EDGE_PIPE 2
#endif
constexpr int kEdgePipe = FESOMTRA_EDGE_PIPE;  // levels of loads in flight in the tiled edge kernel
static_assert(kEdgePipe == 1 || kEdgePipe == 2, "edge pipeline depth 1 or 2");
#ifndef FESOMTRA_EDGE_PAIR
#define FESOMTRA_EDGE_PAIR 1
#endif
// With two levels in flight, stage both in LDS behind one barrier (four LDS buffers).
constexpr bool kEdgePair = FESOMTRA_EDGE_PAIR != 0;
#ifndef FESOMTRA_EDGE_GROUP
#define FESOMTRA_EDGE_GROUP 2
#endif
// Levels staged behind one barrier in the paired form (2 or 4); the next group's loads are in
// flight while a group is computed (2 * group LDS buffers).
constexpr int kEdgeGroup = FESOMTRA_EDGE_GROUP;
static_assert(kEdgeGroup == 2 || kEdgeGroup == 4, "edge level group of 2 or 4");

#ifndef FESOMTRA_EDGE_UNTILED_PF
#define FESOMTRA_EDGE_UNTILED_PF 1
#endif
// Untiled form with one column per wavefront: the node values and the transport of the next
// level are loaded while a level is computed (two register sets).
template <bool TILED, int KB>
constexpr bool kEdgeUntiledPf = FESOMTRA_EDGE_UNTILED_PF != 0 && !TILED && KB == 64;
// Node values of one level for the untiled prefetch: valuesAB at n1, n2, the up-wind triangle's
// and the down-wind triangle's corners; values at n1, n2; the transport.
template <typename C>
struct EdgePf {
    C a[8], t[2], q;
};
template <typename C>
struct ArrLd {
    const C *v;
    __device__ __forceinline__ C operator()(size_t i) const { return v[i]; }
};

// f(std::integral_constant<int, r>) for r = 0 .. N-1 (register sets addressed statically).
template <int N, typename F, int... R>
__device__ __forceinline__ void each_of_impl(F &&f, std::integer_sequence<int, R...>) {
    (f(std::integral_constant<int, R>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void each_of(F &&f) {
    each_of_impl<N>(f, std::make_integer_sequence<int, N>{});
}

// SAME (tiled form only): valuesAB is values, so only one field is staged.  GEN: ensembles (each
// lane reads its member's transport) and cavity meshes.
template <bool TILED, int HOR, bool FCT, int KB, typename S, typename M, typename C, bool SAME = false, bool GEN = false>
__global__ __launch_bounds__(256, (kEdgeWaves<TILED, KB>)) void k_edge(const MeshCol<M> m, const StepPtrs<S, C> p, C num_ord, int lc) {
    // two levels in flight only where the registers allow it (one staged field); the paired
    // form stages a group of G levels behind one barrier
    constexpr bool PAIR = TILED && (SAME ? kEdgePipe : 1) == 2 && kEdgePair;
    constexpr int G = PAIR ? kEdgeGroup : 1;
    constexpr int P = PAIR ? G : (SAME ? kEdgePipe : 1);
    constexpr int NB = PAIR ? 2 * G : 2;  // LDS buffers of the staged field, level nz in nz % NB
    __shared__ C sT[TILED ? NB : 1][TILED ? kHtSlots : 1], sA[TILED && !SAME ? 2 : 1][TILED && !SAME ? kHtSlots : 1];
    EdgeTile<KB, M> et;
    int e, lane, nz0, nz1, lv0, lv1;
    bool act = true;
    if constexpr (TILED) {
        if (!et.init(m, lc)) return;
        e = et.e0 + (int)threadIdx.x / KB;
        lane = (int)threadIdx.x % KB;
        act = e < m.E && lane_on(p.mask, lane);
        if (!act) e = 0;  // in-bounds setup loads; nothing is computed or stored
    } else {
        column_lane<KB>(e, lane);
        if (e >= m.E || !lane_on(p.mask, lane)) return;
    }
    const Member mb = lane_member<GEN>(m, p, lane);
    const int4 q = cld(m.edge_ne, e);
    const int n1 = q.x, n2 = q.y, el1 = q.z, el2 = q.w;
    const int nl1 = cld(m.nlev_e, el1) - 1;
    const int nl2 = el2 >= 0 ? cld(m.nlev_e, el2) - 1 : 0;
    // the edge's levels: from the shallower element top (below a cavity roof) to the deeper bottom
    const int etop = el2 >= 0 ? min(elem_top(m, el1), elem_top(m, el2)) : elem_top(m, el1);
    if constexpr (TILED) {
        nz0 = max(et.nz_b, etop);
        nz1 = min(et.nz_e, max(nl1, nl2));
        act = act && nz0 <= nz1;
        lv0 = et.nz_b;
        lv1 = et.nz_e;
    } else {
        if (!chunk_layers(max(nl1, nl2), lc, nz0, nz1)) return;
        nz0 = max(nz0, etop);
        if (nz0 > nz1) return;
        lv0 = nz0;
        lv1 = nz1;
    }
    const size_t Np = m.Np, Ep = m.Ep;
    const LaneLd<KB, C, S> ldT{p.T, lane}, ldA{p.TAB, lane};
    const bool same = TILED ? SAME : p.TAB == p.T;
    C dx = 0, dy = 0, a = 0;
    int shared_beg = 1, shared_end = 0, u[3] = {0, 0, 0}, d[3] = {0, 0, 0}, nb1 = 0, nb2 = 0, mgb1 = 0, mgb2 = 0;
    C gu[6], gd[6];
    if (HOR != HOR_UPW1) {
        dx = (C)cld(m.edxdy, 2 * (size_t)e);
        dy = (C)cld(m.edxdy, 2 * (size_t)e + 1);
        a = (C)kREarth * (C)cld(m.elem_cos, el1);
        if (el2 >= 0) a = C(0.5) * (a + (C)kREarth * (C)cld(m.elem_cos, el2));
        const int2 ud = cld(m.updn, e);
        if (ud.x >= 0 && ud.y >= 0) {
            // levels where every element around both ends is wet take the up/down-wind
            // triangles' gradients (lib/oce_muscl_adv.F90:414-441)
            shared_end = min(cld(m.nlev_min, n1), cld(m.nlev_min, n2)) - 1;
            if (m.ulev_max) shared_beg = max(cld(m.ulev_max, n1), cld(m.ulev_max, n2));
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                u[c] = cld(m.enodes, (size_t)ud.x * 3 + c);
                d[c] = cld(m.enodes, (size_t)ud.y * 3 + c);
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                gu[c] = (C)cld(m.gsca, (size_t)ud.x * 6 + c);
                gd[c] = (C)cld(m.gsca, (size_t)ud.y * 6 + c);
            }
        }
        if (HOR == HOR_MUSCL) {
            nb1 = cld(m.nbl, n1);
            nb2 = cld(m.nbl, n2);
        }
        if (p.mg) {
            mgb1 = cld(m.mg_base, n1);
            mgb2 = cld(m.mg_base, n2);
        }
    }
    // tile slots of n1, n2, u0..u2, d0..d2 (two 16-bit slots per word)
    int sl[4] = {0, 0, 0, 0};
    if constexpr (TILED) {
#pragma unroll
        for (int i = 0; i < 4; ++i) sl[i] = cld(m.et_loc, 4 * (size_t)e + i);
    }
    auto slot = [&](int k) { return (sl[k >> 1] >> (16 * (k & 1))) & 0xffff; };
    // staging registers and this edge's transport of the next P levels (loads of two levels in
    // flight while a level is computed)
    C stT[P][2], stA[SAME ? 1 : P][2], qn[P];
    auto fetch = [&](int nz, auto set) {
        constexpr int r = decltype(set)::value;
        const size_t o = (size_t)(nz - 1) * Np;
        et.each([&](int i, int node) {
            stT[r][i] = ldT(o + node);
            if constexpr (!SAME)
                if (!same) stA[r][i] = ldA(o + node);
        });
        qn[r] = (act && nz <= nz1) ? (C)cld(m.q, mb.at((size_t)(nz - 1) * Ep + e)) : C(0);
    };
    if constexpr (PAIR) {
        each_of<G>([&](auto r) {
            if (lv0 + r <= lv1) fetch(lv0 + r, r);
        });
    } else if constexpr (TILED) {
        fetch(lv0, std::integral_constant<int, 0>{});
        if (P > 1 && lv0 < lv1) fetch(lv0 + 1, std::integral_constant<int, P - 1>{});
    }
    auto stage = [&](int nz, auto set) {
        constexpr int r = decltype(set)::value;
        et.each([&](int i, int) {
            sT[nz % NB][threadIdx.x + 256 * i] = stT[r][i];
            if constexpr (!SAME)
                if (!same) sA[nz & 1][threadIdx.x + 256 * i] = stA[r][i];
        });
    };
    constexpr bool UPF = kEdgeUntiledPf<TILED, KB>;
    auto pf_load = [&](int nz, EdgePf<C> &f) {
        const size_t o = (size_t)(nz - 1) * Np;
        f.q = (C)cld(m.q, mb.at((size_t)(nz - 1) * Ep + e));
        f.a[0] = ldA(o + n1);
        f.a[1] = ldA(o + n2);
        f.t[0] = f.a[0];
        f.t[1] = f.a[1];
        if (FCT && !same) {
            f.t[0] = ldT(o + n1);
            f.t[1] = ldT(o + n2);
        }
        if (HOR != HOR_UPW1 && nz >= shared_beg && nz <= shared_end) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                f.a[2 + c] = ldA(o + u[c]);
                f.a[5 + c] = ldA(o + d[c]);
            }
        }
    };
    // the edge flux of level nz (LDS: from the values staged for the tiled level loop, qcur its
    // transport; otherwise from memory, or pf: the level's prefetched values in the untiled
    // prefetch form).  The tiled level loop leaves the node-mean levels to a pass after it.
    auto compute = [&](int nz, C qcur, const EdgePf<C> *pf, auto lds_tag) {
        constexpr bool LDS = decltype(lds_tag)::value;
        if (TILED && (!act || nz > nz1 || nz < nz0)) return;
        const bool tri = HOR == HOR_UPW1 || (nz >= shared_beg && nz <= shared_end);
        if (LDS && !tri) return;
        const int buf = nz % NB;
        const int k = nz - 1;
        const C v = LDS ? qcur : (UPF ? pf->q : (C)cld(m.q, mb.at((size_t)k * Ep + e)));
        const size_t o = (size_t)k * Np;
        const size_t ie = ((size_t)k * Ep + e) * KB + lane;
        const SlotLd<KB, C> lsT{sT[TILED ? buf : 0], lane},
            lsA{same ? sT[TILED ? buf : 0] : sA[TILED && !SAME ? (nz & 1) : 0], lane};
        C t1 = 0, t2 = 0;
        if (FCT || same) {
            t1 = LDS ? lsT(slot(0)) : (UPF ? pf->t[0] : ldT(o + n1));
            t2 = LDS ? lsT(slot(1)) : (UPF ? pf->t[1] : ldT(o + n2));
        }
        const C a1 = same ? t1 : (LDS ? lsA(slot(0)) : (UPF ? pf->a[0] : ldA(o + n1)));
        const C a2 = same ? t2 : (LDS ? lsA(slot(1)) : (UPF ? pf->a[1] : ldA(o + n2)));
        C gr[4] = {0, 0, 0, 0}, c1 = 1, c2 = 1;
        if (HOR != HOR_UPW1) {
            if (tri) {
                if constexpr (LDS) {
                    gr[0] = tr_xy<C>(gu, lsA, 0, slot(2), slot(3), slot(4));
                    gr[1] = tr_xy<C>(gd, lsA, 0, slot(5), slot(6), slot(7));
                    gr[2] = tr_xy<C>(gu + 3, lsA, 0, slot(2), slot(3), slot(4));
                    gr[3] = tr_xy<C>(gd + 3, lsA, 0, slot(5), slot(6), slot(7));
                } else if constexpr (UPF) {
                    const ArrLd<C> la{pf->a};
                    gr[0] = tr_xy<C>(gu, la, 0, 2, 3, 4);
                    gr[1] = tr_xy<C>(gd, la, 0, 5, 6, 7);
                    gr[2] = tr_xy<C>(gu + 3, la, 0, 2, 3, 4);
                    gr[3] = tr_xy<C>(gd + 3, la, 0, 5, 6, 7);
                } else if constexpr (!TILED) {  // (the tiled form reads its triangles from LDS only)
                    gr[0] = tr_xy<C>(gu, ldA, o, u[0], u[1], u[2]);
                    gr[1] = tr_xy<C>(gd, ldA, o, d[0], d[1], d[2]);
                    gr[2] = tr_xy<C>(gu + 3, ldA, o, u[0], u[1], u[2]);
                    gr[3] = tr_xy<C>(gd + 3, ldA, o, d[0], d[1], d[2]);
                }
            } else if (p.mg) {  // evaluated by k_mgrad for this pass
                const size_t i1 = (size_t)(mgb1 + nz) * 2 * KB + lane, i2 = (size_t)(mgb2 + nz) * 2 * KB + lane;
                gr[0] = p.mg[i1];
                gr[2] = p.mg[i1 + KB];
                gr[1] = p.mg[i2];
                gr[3] = p.mg[i2 + KB];
            } else {
                node_mean_grad<C>(m, ldA, n1, nz, gr[0], gr[2]);
                node_mean_grad<C>(m, ldA, n2, nz, gr[1], gr[3]);
            }
            if (HOR == HOR_MUSCL) {
                c1 = (nb1 - nz >= 0) ? C(1) : C(0);
                c2 = (nb2 - nz >= 0) ? C(1) : C(0);
            }
        }
        C f;
        if constexpr (std::is_same<C, double>::value) {  // the reference's operation order
            const C fin = FCT ? upw1_edge(t1, t2, v) : C(0);
            if (HOR == HOR_UPW1) f = upw1_edge(a1, a2, v) - fin;
            else f = ho_edge<HOR, C>(a1, a2, gr, v, dx, dy, a, num_ord, c1, c2, fin);
        } else {
            // high-order minus upwind of valuesAB in difference form; with the limiter, plus the
            // upwind difference valuesAB - values.  The upwind fluxes themselves are formed in
            // double where they are summed (k_lo, k_final).
            f = (HOR == HOR_UPW1) ? C(0) : ho_minus_lo<HOR, C>(a1, a2, gr, v, dx, dy, a, num_ord, c1, c2);
            if (FCT && !same) f = f - C(0.5) * ((v + fabs(v)) * (a1 - t1) + (v - fabs(v)) * (a2 - t2));
        }
        p.adfh[ie] = f;
    };
    // one level: stage it, barrier, issue the loads of level nz + P, compute
    auto level = [&](int nz, auto set) {
        constexpr int r = decltype(set)::value;
        C qcur = 0;
        if constexpr (TILED) {
            stage(nz, set);
            qcur = qn[r];
            __syncthreads();
            if (nz + P <= lv1) fetch(nz + P, set);
        }
        compute(nz, qcur, nullptr, std::integral_constant<bool, TILED>{});
    };
    if constexpr (PAIR) {
        // levels nz .. nz + G - 1 staged together behind one barrier; the buffers of the group
        // before are free once every thread has passed it
        for (int nz = lv0; nz <= lv1; nz += G) {
            C qc[G];
            each_of<G>([&](auto r) {
                if (nz + r <= lv1) stage(nz + r, r);
                qc[r] = qn[r];
            });
            __syncthreads();
            each_of<G>([&](auto r) {
                if (nz + G + r <= lv1) fetch(nz + G + r, r);
            });
            each_of<G>([&](auto r) {
                if (nz + r <= lv1) compute(nz + r, qc[r], nullptr, std::true_type{});
            });
        }
    } else if constexpr (UPF) {
        // two register sets, alternating: level nz + 1 loads while level nz computes
        EdgePf<C> f0, f1;
        pf_load(lv0, f0);
        for (int nz = lv0; nz <= lv1; nz += 2) {
            const bool two = nz + 1 <= lv1;
            if (two) pf_load(nz + 1, f1);
            compute(nz, C(0), &f0, std::false_type{});
            if (two) {
                if (nz + 2 <= lv1) pf_load(nz + 2, f0);
                compute(nz + 1, C(0), &f1, std::false_type{});
            }
        }
    } else {
        // unrolled by the pipeline depth, so every register set is addressed statically
        for (int nz = lv0; nz <= lv1; nz += P) {
            level(nz, std::integral_constant<int, 0>{});
            if constexpr (P > 1)
                if (nz + 1 <= lv1) level(nz + 1, std::integral_constant<int, P - 1>{});
        }
    }
    if constexpr (TILED && HOR != HOR_UPW1) {
        // the node-mean levels (outside the triangles' shared wet range), from memory and free of
        // the tile's barriers: the top levels below a cavity roof and the bottom levels
        if (act) {
            const std::false_type direct{};
            if (shared_end < shared_beg) {
                for (int nz = nz0; nz <= nz1; ++nz) compute(nz, C(0), nullptr, direct);
            } else {
                for (int nz = nz0; nz <= min(nz1, shared_beg - 1); ++nz) compute(nz, C(0), nullptr, direct);
                for (int nz = max(nz0, shared_end + 1); nz <= nz1; ++nz) compute(nz, C(0), nullptr, direct);
            }
        }
    }
}

// Node-mean gradients of valuesAB for the entries MeshCol::mg_* lists (lib/oce_muscl_adv.F90:394-409,
// the same evaluation as node_mean_grad in the edge kernel, so the stored values are those it
// would compute): one thread per entry and lane.  Separate from the edge kernel, whose tiles
// would otherwise wait at their next barrier for these few serial gathers.
template <int KB, typename S, typename M, typename C>
__global__ __launch_bounds__(256) void k_mgrad(const MeshCol<M> m, const StepPtrs<S, C> p) {
    int i, lane;
    column_lane<KB>(i, lane);
    if (i >= m.mg_n || !lane_on(p.mask, lane)) return;
    const int n = cld(m.mg_node, i), nz = cld(m.mg_lev, i);
    const LaneLd<KB, C, S> ldA{p.TAB, lane};
    C gx, gy;
    node_mean_grad<C>(m, ldA, n, nz, gx, gy);
    p.mg[(size_t)i * 2 * KB + lane] = gx;
    p.mg[((size_t)i * 2 + 1) * KB + lane] = gy;
}

// Edge transport of the current velocity, q(k, e) = sum over the edge's elements wet at level k
// of (+-) (u dy - v dx) helem (lib/oce_adv_tra_hor.F90:166-253, segments A to E: below a cavity
// roof only the element already wet contributes), in double, once per step for all tracers.  One
// thread per edge walks its levels; levels where neither element is wet are not read.  In an
// ensemble (ENS) consecutive threads take the members of one edge: each reads its member's uv
// and a wavefront stores contiguous runs of the member-interleaved q.
template <typename M, bool ENS = false>
__global__ __launch_bounds__(256) void k_qflux(const MeshCol<M> m) {
    int e = column_index();
    Member mb{0, 1};
    if constexpr (ENS) {
        mb = Member{e % m.nmem, m.nmem};
        e /= m.nmem;
    }
    if (e >= m.E) return;
    const M *__restrict__ uv = ENS ? m.uv + (size_t)mb.m * (size_t)m.L * m.Tp * 2 : m.uv;
    const int4 q = cld(m.edge_ne, e);
    const int el1 = q.z, el2 = q.w;
    const int nl1 = cld(m.nlev_e, el1) - 1;
    const int nl2 = el2 >= 0 ? cld(m.nlev_e, el2) - 1 : 0;
    const int nu1 = elem_top(m, el1), nu2 = el2 >= 0 ? elem_top(m, el2) : 1;
    const double cd0 = cld(m.ecd, 4 * (size_t)e), cd1 = cld(m.ecd, 4 * (size_t)e + 1);
    const double cd2 = cld(m.ecd, 4 * (size_t)e + 2), cd3 = cld(m.ecd, 4 * (size_t)e + 3);
    const size_t Tp = m.Tp, Ep = m.Ep;
    const int top = max(nl1, nl2);
    // Levels are independent: the loads are unconditional (a dry element's level is clamped to
    // a wet one, a missing element replaced by the first, and the value not used) and the loop
    // is unrolled, so the loads of several levels are in flight together.
    const int e2 = el2 >= 0 ? el2 : el1, w1 = max(nl1, 1), w2 = max(nl2, 1);
    auto level = [&](int nz, double u1, double y1, double h1, double u2, double y2, double h2) {
        const bool use1 = nz <= nl1 && nz >= nu1, use2 = nz <= nl2 && nz >= nu2;
        const double v1 = (-y1 * cd0 + u1 * cd1) * h1;
        const double v2 = (y2 * cd2 - u2 * cd3) * h2;
        m.q[mb.at((size_t)(nz - 1) * Ep + e)] = (use1 && use2) ? v1 + v2 : (use1 ? v1 : v2);
    };
    auto idx = [&](int nz, int w, int el) { return (size_t)(min(nz, w) - 1) * Tp + el; };
    constexpr int B = 4;  // levels per batch: all their loads are issued before the first store
    if (top >= B) {
        for (int nb = 1; nb <= top; nb += B) {
            const int n0 = min(nb, top - B + 1);  // the last batch moves back and rewrites equal values
            double u1[B], y1[B], h1[B], u2[B], y2[B], h2[B];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const size_t i1 = idx(n0 + j, w1, el1), i2 = idx(n0 + j, w2, e2);
                u1[j] = (double)cld(uv, 2 * i1);
                y1[j] = (double)cld(uv, 2 * i1 + 1);
                h1[j] = (double)cld(m.helem, i1);
                u2[j] = (double)cld(uv, 2 * i2);
                y2[j] = (double)cld(uv, 2 * i2 + 1);
                h2[j] = (double)cld(m.helem, i2);
            }
#pragma unroll
            for (int j = 0; j < B; ++j) level(n0 + j, u1[j], y1[j], h1[j], u2[j], y2[j], h2[j]);
        }
    } else {
        for (int nz = 1; nz <= top; ++nz) {
            const size_t i1 = idx(nz, w1, el1), i2 = idx(nz, w2, e2);
            level(nz, (double)cld(uv, 2 * i1), (double)cld(uv, 2 * i1 + 1), (double)cld(m.helem, i1),
                  (double)cld(uv, 2 * i2), (double)cld(uv, 2 * i2 + 1), (double)cld(m.helem, i2));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Node -> edge entries cached in registers (the first NC, the rest read from memory).
// The far end is a node number (m.ne_other) or, with PACK (halo tiles), the far node's tile
// slot (m.ht_loc, < 2^15) kept with the entry's deepest level in one register.
// Six entries cover an interior node of a regular triangulation and most nodes of an ocean mesh;
// eight spilled registers in the tiled FP64 node kernels (1.25 -> 1.12 ms for the low-order
// kernel on the 1000x1000x48 mesh with six).
#ifndef FESOMTRA_DEG_CACHE_NODE
#define FESOMTRA_DEG_CACHE_NODE 6
#endif
#ifndef FESOMTRA_DEG_CACHE_WIDE
#define FESOMTRA_DEG_CACHE_WIDE 6
#endif
// An entry's level range (the edge's scatter range) is kept as bottom | top << 8 (levels < 256,
// checked at setup); a slot past the node's degree holds 0, so it is active at no level (no
// separate test of the degree, whose per-slot masks would take scalar registers).  GEN: kernels for meshes with cavities (the top is 1 otherwise, and the
// other kernels test the bottom only).
template <bool PACK, int NC, bool GEN = false>
struct NodeEdges {
    static constexpr int kN = NC;
    int b, deg;
    int e2[NC], top[NC], oth[PACK ? 1 : NC];
    __device__ __forceinline__ int top_of(int i) const { return top[i] & 0xff; }
    __device__ __forceinline__ int utop_of(int i) const { return (top[i] >> 8) & 0xff; }
    __device__ __forceinline__ bool on(int i, int nz) const {
        if constexpr (GEN) return nz <= top_of(i) && nz >= utop_of(i);
        else return nz <= top_of(i);
    }
    __device__ __forceinline__ int far_of(int i) const { return PACK ? (top[i] >> 16) : oth[i]; }
};
template <typename M>
__device__ __forceinline__ int entry_utop(const MeshCol<M> &m, int i) {
    return m.ne_utop ? cld(m.ne_utop, i) : 1;
}
// Entry i of the CSR lists is active at level nz.
template <bool GEN, typename M>
__device__ __forceinline__ bool entry_on(const MeshCol<M> &m, int i, int nz) {
    if constexpr (GEN) return nz <= cld(m.ne_top, i) && nz >= entry_utop(m, i);
    else return nz <= cld(m.ne_top, i);
}
template <bool PACK, int NC, bool GEN, typename M>
__device__ __forceinline__ void load_edges(const MeshCol<M> &m, int n, NodeEdges<PACK, NC, GEN> &ne) {
    ne.b = cld(m.ne_ptr, n);
    ne.deg = cld(m.ne_ptr, n + 1) - ne.b;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const bool on = i < ne.deg;
        ne.e2[i] = on ? cld(m.ne_e2, ne.b + i) : 0;
        const int lv = on ? cld(m.ne_top, ne.b + i) | ((GEN ? entry_utop(m, ne.b + i) : 1) << 8) : 0;
        if constexpr (PACK) {
            ne.top[i] = on ? (cld(m.ht_loc, ne.b + i) << 16) | lv : 0;
        } else {
            ne.top[i] = lv;
            ne.oth[i] = on ? cld(m.ne_other, ne.b + i) : 0;
        }
    }
}
template <bool PACK, int NC, bool GEN>
__device__ __forceinline__ void no_edges(NodeEdges<PACK, NC, GEN> &ne) {
    ne.b = ne.deg = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) ne.e2[i] = ne.top[i] = 0;
#pragma unroll
    for (int i = 0; i < (PACK ? 1 : NC); ++i) ne.oth[i] = 0;
}
// One column per wavefront (KB = 64): the cached entries live in scalar registers.  Marking them
// as redefined at the top of every level keeps the compiler from hoisting each entry's 64-bit
// addresses out of the level loop: held across the loop they exceed the scalar register file and
// are spilled to vector lanes and read back at every level.
template <bool PACK, int NC, bool GEN>
__device__ __forceinline__ void scalar_fence(NodeEdges<PACK, NC, GEN> &ne) {
#pragma unroll
    for (int i = 0; i < NC; ++i) asm volatile("" : "+s"(ne.e2[i]));
#pragma unroll
    for (int i = 0; i < (PACK ? 1 : NC); ++i) asm volatile("" : "+s"(ne.oth[i]));
}
// Calls f(edge, second, far) for the edges of n active at level nz, in reference order.
template <bool PACK, int NC, bool GEN, typename M, typename F>
__device__ __forceinline__ void for_edges(const MeshCol<M> &m, const NodeEdges<PACK, NC, GEN> &ne, int nz, F &&f) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
        if (ne.on(i, nz)) f(ne.e2[i] >> 1, (ne.e2[i] & 1) != 0, ne.far_of(i));
    for (int i = NC; i < ne.deg; ++i) {
        const int e2 = cld(m.ne_e2, ne.b + i);
        if (entry_on<GEN>(m, ne.b + i, nz))
            f(e2 >> 1, (e2 & 1) != 0, cld(PACK ? m.ht_loc : m.ne_other, ne.b + i));
    }
}

// Two-phase form of for_edges: load(edge, second, far) for every edge active at level nz is
// issued first (the loads of all cached entries are in flight together), then use(value, edge,
// second, far) consumes them in reference order, so a node's sum waits for one memory latency
// per level instead of one per edge.  It holds all loaded values at once: where that would spill
// (the untiled low-order kernel of narrow blocks, whose columns hold more per-thread state) the
// one-phase form is kept (TWO = false).
template <int KB>
constexpr bool kTwoPhase = KB > 4;
// The flux divergence divides every edge flux of a cell by the cell's volume; with one column
// per wavefront (KB = 64, bound by vector instructions) it does so through one reciprocal per
// cell (div_reuse, bit-identical to dividing).
template <int KB>
constexpr bool kDivReuse = KB == 64;
// Node -> edge entries held in registers: with one column per wavefront (KB = 64) they are
// scalar registers, which the wide kernels are short of.
template <int KB>
constexpr int kDegCacheFor = KB == 64 ? FESOMTRA_DEG_CACHE_WIDE : FESOMTRA_DEG_CACHE_NODE;

// The load phase of the cached entries active at level nz, into v; edges_use consumes them
// (and loads and consumes the entries past the cache).  Split so that other loads can be issued
// between the two.
template <bool PACK, int NC, bool GEN, typename LOAD, typename V>
__device__ __forceinline__ void edges_pre(const NodeEdges<PACK, NC, GEN> &ne, int nz, LOAD &&load, V (&v)[NC]) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
        if (ne.on(i, nz)) v[i] = load(ne.e2[i] >> 1, (ne.e2[i] & 1) != 0, ne.far_of(i));
}
// As edges_pre, but loading every cached slot whatever the level (no branch around a load, so
// the loads are issued back to back); a slot past the node's degree names edge 0 and node 0.
// For loads whose index stays in bounds for any cached entry at any level of the column.
template <bool PACK, int NC, bool GEN, typename LOAD, typename V>
__device__ __forceinline__ void edges_pre_all(const NodeEdges<PACK, NC, GEN> &ne, LOAD &&load, V (&v)[NC]) {
#pragma unroll
    for (int i = 0; i < NC; ++i) v[i] = load(ne.e2[i] >> 1, (ne.e2[i] & 1) != 0, ne.far_of(i));
}
template <bool PACK, int NC, bool GEN, typename M, typename V, typename LOAD, typename USE>
__device__ __forceinline__ void edges_use(const MeshCol<M> &m, const NodeEdges<PACK, NC, GEN> &ne, int nz, const V (&v)[NC],
                                          LOAD &&load, USE &&use) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
        if (ne.on(i, nz)) use(v[i], ne.e2[i] >> 1, (ne.e2[i] & 1) != 0, ne.far_of(i));
    for (int i = NC; i < ne.deg; ++i) {
        const int e2 = cld(m.ne_e2, ne.b + i);
        if (entry_on<GEN>(m, ne.b + i, nz)) {
            const int far = cld(PACK ? m.ht_loc : m.ne_other, ne.b + i);
            use(load(e2 >> 1, (e2 & 1) != 0, far), e2 >> 1, (e2 & 1) != 0, far);
        }
    }
}
template <bool TWO, bool PACK, int NC, bool GEN, typename M, typename LOAD, typename USE>
__device__ __forceinline__ void for_edges_ld(const MeshCol<M> &m, const NodeEdges<PACK, NC, GEN> &ne, int nz, LOAD &&load,
                                             USE &&use) {
    if constexpr (!TWO) {
        for_edges(m, ne, nz, [&](int e, bool second, int far) { use(load(e, second, far), e, second, far); });
        return;
    }
    using V = decltype(load(0, false, 0));
    V v[NC];
    edges_pre(ne, nz, load, v);
    edges_use(m, ne, nz, v, load, use);
}

// Low-order solution and vertical antidiffusive flux (lib/oce_adv_tra_driver.F90:102-216 and
// the high-order vertical call :266-282 with o_init_zero=.false., pwvel = w).  Interface fluxes
// are evaluated top-down in double; layer nz uses the fluxes at its top (nz) and bottom (nz + 1).
// The upwind edge fluxes are formed here in double from q and the two end values (never stored);
// TILED stages the neighbours' values per level in LDS.  The low-order vertical flux uses w_e;
// with wsplit (:221-231) the flux subtracted from the high-order one is the upwind flux of the
// full w instead, and k_vimpl then adds the implicit part to the stored LO.
template <bool TILED, int VER, int KB, typename S, typename M, typename C, bool GEN = false>
__global__ __launch_bounds__(256, (kWaves<TILED, KB>)) void k_lo(const MeshCol<M> m, const StepPtrs<S, C> p, C num_ord, double dt_step, int lc,
                                                                 bool wsplit) {
    using D = double;
    __shared__ D sT[TILED ? 2 : 1][TILED ? kHtSlots : 1];
    // tiled form with one velocity member: the transports of the tile's edges are staged per
    // level too (QS per thread), fetched a level ahead with the neighbour values
    constexpr bool QST = TILED && !GEN;
    constexpr int QS = KB == 1 ? 4 : 2;
    __shared__ double sQ[QST ? 2 : 1][QST ? QS * 256 : 1];
    HaloTile<KB, M> ht;
    int n, lane, nzmax, L, top, nz0, nz1, lv0, lv1;
    bool act = true;
    if constexpr (TILED) {
        if (!ht.init(m, lc)) return;
        n = ht.n0 + (int)threadIdx.x / KB;
        lane = (int)threadIdx.x % KB;
        act = n < m.N && lane_on(p.mask, lane);
        nzmax = act ? cld(m.nlev_n, n) : 1;
        top = act ? node_top<GEN>(m, n) : 1;
        L = nzmax - 1;
        nz0 = max(ht.nz_b, top);
        nz1 = min(ht.nz_e, L);
        act = act && nz0 <= nz1;
        lv0 = ht.nz_b;
        lv1 = ht.nz_e;
    } else {
        column_lane<KB>(n, lane);
        if (n >= m.N || !lane_on(p.mask, lane)) return;
        nzmax = cld(m.nlev_n, n);
        top = node_top<GEN>(m, n);
        L = nzmax - 1;
        if (!chunk_layers(L, lc, nz0, nz1)) return;
        nz0 = max(nz0, top);
        if (nz0 > nz1) return;
        lv0 = nz0;
        lv1 = nz1;
    }
    const size_t Np = m.Np, Ep = m.Ep;
    NodeEdges<TILED, kDegCacheFor<KB>, GEN> ne;
    if (act) load_edges(m, n, ne);
    else no_edges(ne);
    const int nc_n = act ? n : 0;
    const double dt = node_dt<GEN>(m, nc_n, dt_step);
    const Member mb = lane_member<GEN>(m, p, lane);
    const NodeCol<D, S, M, KB> nc{Np, nc_n, lane, L, mb};
    const Col<D, S, M, KB> col = make_col<D, S, M, KB>(p.TAB, m.w, m, nc_n, lane, nzmax, top, mb);
    const D pv = (D)num_ord;
    D tw[4], z[4], lo_top = 0, adf_top = 0, tcur = 0;
    QuotWin<D> qw;
    if (act) {
        const int j = nz0;  // top interface of the chunk
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] = (VER == VER_QR4C) ? nc.mlay(m.Z, j - 2 + i) : D(0);
        const D wej = nc.vel(m.we, j), wj = nc.vel(m.w, j), aj = nc.itf(m.area, j);
        const D zbj = (VER == VER_QR4C) ? nc.itf(m.zbar, j) : D(0);
#pragma unroll
        for (int i = 0; i < 4; ++i) tw[i] = nc.lay(p.TAB, j - 2 + i);
        tcur = nc.lay(p.T, j);
        const D tup = nc.lay(p.T, j - 1);
        lo_top = ver_upw1_w<D>(j, top, nzmax, tup, tcur, wej, aj, D(0));
        const D fin_top = wsplit ? ver_upw1_w<D>(j, top, nzmax, tup, tcur, wj, aj, D(0)) : lo_top;
        if (VER == VER_QR4C) qw.init(tw, z);
        adf_top = ver_ho_at<VER>(j, nzmax, tw, z, zbj, wj, aj, pv, dt, col, fin_top, qw.q);
    }
    // staging registers of the next level (tiled form; the neighbour values of one level in
    // flight while a level is computed: staging two levels behind one barrier measured slower)
    D st[2];
    // staged transports: this thread's edges of the tile (slot j * 256 + thread; a slot past the
    // tile's edges loads edge 0 and is never read), and each cached entry's slot * 2 + second
    int qe[QS], ce[decltype(ne)::kN];
    double sq[QS];
    if constexpr (QST) {  // the host builds tiles only where every tile's edges fit the slots
        const int t = ht.n0 / HaloTile<KB, M>::TC, eb = cld(m.ht_eptr, t), ec = cld(m.ht_eptr, t + 1) - eb;
#pragma unroll
        for (int j = 0; j < QS; ++j) {
            const int sl = j * 256 + (int)threadIdx.x;
            qe[j] = sl < ec ? cld(m.ht_edges, eb + sl) : 0;
        }
#pragma unroll
        for (int i = 0; i < decltype(ne)::kN; ++i) ce[i] = (act && i < ne.deg) ? cld(m.ht_eloc, ne.b + i) : 0;
    }
    auto fetch = [&](int nz) {
        ht.each([&](int i, int node) { st[i] = ldc<D>(p.T, ((size_t)(nz - 1) * Np + node) * KB + lane); });
        if constexpr (QST) {
#pragma unroll
            for (int j = 0; j < QS; ++j) sq[j] = (double)cld(m.q, (size_t)(nz - 1) * Ep + qe[j]);
        }
    };
    auto stage = [&](int nz) {
        ht.each([&](int i, int) { sT[nz & 1][threadIdx.x + 256 * i] = st[i]; });
        if constexpr (QST) {
#pragma unroll
            for (int j = 0; j < QS; ++j) sQ[nz & 1][j * 256 + threadIdx.x] = sq[j];
        }
    };
    // The per-lane loads of a layer (the window's next value, the next layer's value and, untiled,
    // the far values of the cached edges with their transports) are issued together before any
    // is used: unconditionally, at a level clamped into the column, the value then selected, so
    // the compiler does not wait for each load inside its own branch.
    auto lay_raw = [&](const S *a, int lz) -> S {
        const int c = min(max(lz, 1), max(L, 1));
        return a[((size_t)(c - 1) * Np + n) * KB + lane];
    };
    auto lay_of = [&](S raw, int lz) -> D { return (lz >= 1 && lz <= L) ? ldc<D>(&raw, 0) : D(0); };
    struct EdgeRaw {
        S t;
        double q;
    };
    // layer nz of this column (its neighbours' values staged in sT[nz & 1] when TILED)
    auto body = [&](int nz) {
        if (TILED && (!act || nz > nz1 || (GEN && nz < nz0))) return;
        const int buf = nz & 1;
        const int j = nz + 1;  // bottom interface of layer nz
        const size_t eo = (size_t)(nz - 1) * Ep, no = (size_t)(nz - 1) * Np;
        if constexpr (KB == 64) scalar_fence(ne);
        const S raw3 = lay_raw(p.TAB, j + 1), rawn = lay_raw(p.T, j);
        auto edge_raw = [&](int e, bool, int oth) {
            return EdgeRaw{p.T[(no + oth) * KB + lane], (double)cld(m.q, mb.at(eo + e))};
        };
        EdgeRaw er[TILED ? 1 : decltype(ne)::kN];
        if constexpr (!TILED) edges_pre_all(ne, edge_raw, er);
        if (VER == VER_QR4C) {
            z[0] = z[1];
            z[1] = z[2];
            z[2] = z[3];
            z[3] = nc.mlay(m.Z, j + 1);
        }
        const bool bottom = (j == nzmax);
        const D wej = nc.vel(m.we, j), wj = nc.vel(m.w, j), aj = nc.itf(m.area, j);
        const D zbj = (VER == VER_QR4C && !bottom) ? nc.itf(m.zbar, j) : D(0);
        const size_t cell = (size_t)(nz - 1) * Np + n;
        const double h = (double)cld(m.hnode, cell), hn = (double)cld(m.hnode_new, cell);
        const double V = (double)cld(m.areasvol, cell);
        tw[0] = tw[1];
        tw[1] = tw[2];
        tw[2] = tw[3];
        tw[3] = lay_of(raw3, j + 1);
        if (VER == VER_QR4C) qw.slide(tw, z);
        const D tnext = lay_of(rawn, j);
        const D lo_bot = ver_upw1_w<D>(j, top, nzmax, tcur, tnext, wej, aj, D(0));
        const D fin_bot = wsplit ? ver_upw1_w<D>(j, top, nzmax, tcur, tnext, wj, aj, D(0)) : lo_bot;
        const D adf_bot = ver_ho_at<VER>(j, nzmax, tw, z, zbj, wj, aj, pv, dt, col, fin_bot, qw.q);
        // upwind edge fluxes (lib/oce_adv_tra_hor.F90:178-181) from q and the two end values,
        // summed in reference edge order
        double s = 0.0;
        auto add_edge = [&](double to, double qe, bool second) {
            double f;
            if constexpr (std::is_same<C, double>::value) {
                f = upw1_edge<double>(second ? to : tcur, second ? tcur : to, qe);
            } else {
                // single precision: -0.5 (t1 (v + |v|) + t2 (v - |v|)) is -v t1 for v >= 0 and
                // -v t2 otherwise (one product instead of four operations; both ends of an edge
                // form the same value, so the fluxes stay antisymmetric)
                f = -qe * (qe >= 0.0 ? (second ? to : tcur) : (second ? tcur : to));
            }
            s = second ? s - f : s + f;
        };
        if constexpr (QST) {
            // transports and neighbour values from the tile, in reference edge order
#pragma unroll
            for (int i = 0; i < decltype(ne)::kN; ++i)
                if (ne.on(i, nz)) add_edge(sT[buf][ne.far_of(i) * KB + lane], sQ[buf][ce[i] >> 1], (ce[i] & 1) != 0);
            for (int i = decltype(ne)::kN; i < ne.deg; ++i) {
                if (!entry_on<GEN>(m, ne.b + i, nz)) continue;
                const int c = cld(m.ht_eloc, ne.b + i);
                add_edge(sT[buf][cld(m.ht_loc, ne.b + i) * KB + lane], sQ[buf][c >> 1], (c & 1) != 0);
            }
        } else if constexpr (TILED) {
            // transports from memory, issued together; neighbour values from the tile
            for_edges_ld<true>(
                m, ne, nz, [&](int e, bool, int) { return (double)cld(m.q, mb.at(eo + e)); },
                [&](double qe, int, bool second, int oth) { add_edge(sT[buf][oth * KB + lane], qe, second); });
        } else {
            edges_use(m, ne, nz, er, edge_raw, [&](const EdgeRaw &v, int, bool second, int) {
                add_edge(ldc<double>(&v.t, 0), v.q, second);
            });
        }
        const double lo = ((double)tcur * h + (s + ((double)lo_top - (double)lo_bot)) * dt / V) / hn;
        // FP64 stores LO; single precision stores the increment LO - T, so that low-order changes
        // far below the tracer's own magnitude survive the rounding to C
        p.LO[cell * KB + lane] = std::is_same<C, double>::value ? (C)lo : (C)(lo - (double)tcur);
        if constexpr (!std::is_same<C, double>::value)
            if (p.LOe) {  // compensated storage: what the single-precision increment lost
                const double inc = lo - (double)tcur;
                p.LOe[cell * KB + lane] = (C)(inc - (double)(C)inc);
            }
        p.adfv[cell * KB + lane] = (C)adf_top;
        if (bottom) p.adfv[(cell + Np) * KB + lane] = (C)adf_bot;
        lo_top = lo_bot;
        adf_top = adf_bot;
        tcur = tnext;
    };
    if constexpr (!TILED) {
        for (int nz = lv0; nz <= lv1; ++nz) body(nz);
    } else {
        fetch(lv0);
        for (int nz = lv0; nz <= lv1; ++nz) {
            stage(nz);
            __syncthreads();
            if (nz < lv1) fetch(nz + 1);
            body(nz);
        }
    }
}

// Implicit first-order upwind vertical advection of the low-order solution with the implicit
// part w_i of the split vertical velocity (lib/oce_adv_tra_ver.F90:90-244, called from
// lib/oce_adv_tra_driver.F90:221-226 when dynamics%use_wsplit): per column and lane, the
// tridiagonal system on hnode_new is built and eliminated top-down in double, storing (c', t')
// in p.tri, then back-substituted bottom-up and added to LO.  Operation order as the reference
// (FP64 results are bit-identical).  Columns have at least two wet layers (checked at setup).
template <int KB, typename S, typename M, typename C, bool GEN = false>
__global__ __launch_bounds__(256) void k_vimpl(const MeshCol<M> m, const StepPtrs<S, C> p, double dt_step) {
    int n, lane;
    column_lane<KB>(n, lane);
    if (n >= m.N || !lane_on(p.mask, lane)) return;
    const double dt = node_dt<GEN>(m, n, dt_step);
    const int L = cld(m.nlev_n, n) - 1, top = node_top<GEN>(m, n);
    if (L - top < 1) return;
    const size_t Np = m.Np;
    auto at = [&](int nz) { return ((size_t)(nz - 1) * Np + n) * KB + lane; };
    // LO in double (single precision stores the increment LO - T)
    auto lo_at = [&](int nz) -> double {
        if constexpr (std::is_same<C, double>::value) return p.LO[at(nz)];
        else return ldc<double>(p.T, at(nz)) + ((double)p.LO[at(nz)] + (p.LOe ? (double)p.LOe[at(nz)] : 0.0));
    };
    auto lo_put = [&](int nz, double v) {
        if constexpr (std::is_same<C, double>::value) {
            p.LO[at(nz)] = v;
        } else {
            const double inc = v - ldc<double>(p.T, at(nz));
            p.LO[at(nz)] = (C)inc;
            if (p.LOe) p.LOe[at(nz)] = (C)(inc - (double)(C)inc);
        }
    };
    const double zinv = 1.0 * dt;
    double t0 = 0.0, t1 = lo_at(top), t2 = lo_at(top + 1), cpp = 0.0, tpp = 0.0;
    const size_t c0 = (size_t)(top - 1) * Np + n;  // the column's top interface
    const Member mb = lane_member<GEN>(m, p, lane);
    double w1 = (double)cld(m.wi, mb.at(c0)), a1 = (double)cld(m.area, c0);
    for (int nz = top; nz <= L; ++nz) {
        const size_t cell = (size_t)(nz - 1) * Np + n;
        const double hn = (double)cld(m.hnode_new, cell), V = (double)cld(m.areasvol, cell);
        double a, b, c, v = zinv * a1 / V;
        if (nz == top) {
            a = 0.0;
            b = hn + w1 * v;
        } else {
            a = cmin(0.0, w1) * v;
            b = hn + cmax(0.0, w1) * v;
        }
        double w2 = 0.0, a2 = 0.0;
        if (nz < L) {
            w2 = (double)cld(m.wi, mb.at(cell + Np));
            a2 = (double)cld(m.area, cell + Np);
            v = zinv * a2 / V;
            b = b - cmin(0.0, w2) * v;
            c = -cmax(0.0, w2) * v;
        } else {
            c = 0.0;
        }
        double tr;
        if (nz == top) tr = -(b - hn) * t1 - c * t2;
        else if (nz < L) tr = -a * t0 - (b - hn) * t1 - c * t2;
        else tr = -a * t0 - (b - hn) * t1;
        double cp, tp;
        if (nz == top) {
            cp = c / b;
            tp = tr / b;
        } else {
            const double mm = b - cpp * a;
            cp = c / mm;
            tp = (tr - tpp * a) / mm;
        }
        if (nz < L) p.tri[at(nz)] = make_double2(cp, tp);
        cpp = cp;
        tpp = tp;
        w1 = w2;
        a1 = a2;
        t0 = t1;
        t1 = t2;
        t2 = nz + 2 <= L ? lo_at(nz + 2) : 0.0;
    }
    double x = tpp;
    lo_put(L, t0 + x);
    for (int nz = L - 1; nz >= top; --nz) {
        const double2 ct = p.tri[at(nz)];
        x = ct.y - ct.x * x;
        lo_put(nz, lo_at(nz) + x);
    }
}

// Vertical high-order flux without the limiter (pwvel = we, o_init_zero=.true.), in double:
// interfaces nz0..nz1 of the chunk, plus the bottom interface in the deepest chunk.
template <int VER, int KB, typename S, typename M, typename C, bool GEN = false>
__global__ __launch_bounds__(256, kMinWaves<KB>) void k_vert(const MeshCol<M> m, const StepPtrs<S, C> p, C num_ord, double dt_step, int lc) {
    using D = double;
    int n, lane;
    column_lane<KB>(n, lane);
    if (n >= m.N || !lane_on(p.mask, lane)) return;
    const double dt = node_dt<GEN>(m, n, dt_step);
    const int nzmax = cld(m.nlev_n, n), L = nzmax - 1, top = node_top<GEN>(m, n);
    int nz0, nz1;
    if (!chunk_layers(L, lc, nz0, nz1)) return;
    nz0 = max(nz0, top);
    if (nz0 > nz1) return;
    const size_t Np = m.Np;
    const Member mb = lane_member<GEN>(m, p, lane);
    const NodeCol<D, S, M, KB> nc{Np, n, lane, L, mb};
    const Col<D, S, M, KB> col = make_col<D, S, M, KB>(p.TAB, m.we, m, n, lane, nzmax, top, mb);
    const D pv = (D)num_ord;
    D tw[4], z[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] = (VER == VER_QR4C) ? nc.mlay(m.Z, nz0 - 3 + i) : D(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) tw[i] = nc.lay(p.TAB, nz0 - 3 + i);
    QuotWin<D> qw;  // primed so that the first slide below gives the window of interface nz0
    if (VER == VER_QR4C) {
        qw.q[1] = QuotWin<D>::quot(tw[1], tw[2], z[1], z[2]);
        qw.q[2] = QuotWin<D>::quot(tw[2], tw[3], z[2], z[3]);
    }
    const int jend = nz1 == L ? nzmax : nz1;
    for (int j = nz0; j <= jend; ++j) {
        if (VER == VER_QR4C) {
            z[0] = z[1];
            z[1] = z[2];
            z[2] = z[3];
            z[3] = nc.mlay(m.Z, j + 1);
        }
        const bool bottom = (j == nzmax);
        const D wj = nc.vel(m.we, j), aj = nc.itf(m.area, j);
        const D zbj = (VER == VER_QR4C && !bottom) ? nc.itf(m.zbar, j) : D(0);
        tw[0] = tw[1];
        tw[1] = tw[2];
        tw[2] = tw[3];
        tw[3] = nc.lay(p.TAB, j + 1);
        if (VER == VER_QR4C) qw.slide(tw, z);
        p.adfv[((size_t)(j - 1) * Np + n) * KB + lane] =
            (C)ver_ho_at<VER>(j, nzmax, tw, z, zbj, wj, aj, pv, dt, col, D(0), qw.q);
    }
}

// ---------------------------------------------------------------------------------------------
// Cavity meshes: the values the limiter's cluster bound finds above an element's top.  The
// reference keeps the element bounds in the edge-gradient work array (lib/oce_adv_tra_fct.F90:77,
// 129-147) and writes them only from the element's top down, while the cluster bound of a node
// (:164-184) reads every element around it at every level of the node: above the element's top it
// finds the gradient fill (lib/oce_muscl_adv.F90:361-530) of the edge that has the element's
// number, components 1 and 2 (the x gradients at the edge's first and second node), where that
// fill wrote the level, and zero elsewhere.  One thread per (slot, lane) forms them from valuesAB
// for the levels above the slot element's top, with the fill's operation order; fill = 0 (a pass
// whose gradients were never filled) stores zeros.
template <int KB, typename S, typename M, typename C>
__global__ __launch_bounds__(256) void k_cavity_aux(const MeshCol<M> m, const StepPtrs<S, C> p, int fill) {
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int sl = g / KB, lane = g % KB;
    if (sl >= m.cq_n || !lane_on(p.mask, lane)) return;
    const int rows = cld(m.cq_top, sl) - 1;
    const size_t off = (size_t)cld(m.cq_off, sl);
    auto put = [&](int nz, int comp, C v) { p.cqv[((off + (size_t)(nz - 1)) * 2 + comp) * KB + lane] = v; };
    if (!fill) {
        for (int nz = 1; nz <= rows; ++nz) {
            put(nz, 0, C(0));
            put(nz, 1, C(0));
        }
        return;
    }
    const int e = cld(m.cq_edge, sl);
    const int4 q = cld(m.edge_ne, e);
    const int2 ud = cld(m.updn, e);
    const bool both = ud.x >= 0 && ud.y >= 0;
    int nzmin = 1, nzmax = 0;  // the levels that take the up/down-wind triangles' gradients
    if (both) {
        nzmin = max(cld(m.ulev_max, q.x), cld(m.ulev_max, q.y));
        nzmax = min(cld(m.nlev_min, q.x), cld(m.nlev_min, q.y));
    }
    const LaneLd<KB, C, S> ldA{p.TAB, lane};
    for (int nz = 1; nz <= rows; ++nz) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int node = side ? q.y : q.x, tri = side ? ud.y : ud.x;
            const int nb = cld(m.nlev_n, node) - 1;
            const int last = both ? max(nzmin - 1, nb) : nb;  // the fill's loops write up to here
            C v = C(0);
            if (nz >= node_top(m, node) && nz <= last) {
                if (both && nz >= nzmin && nz <= nzmax - 1) {
                    C gc[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) gc[c] = (C)cld(m.gsca, (size_t)tri * 6 + c);
                    v = tr_xy<C>(gc, ldA, (size_t)(nz - 1) * m.Np, cld(m.enodes, (size_t)tri * 3),
                                 cld(m.enodes, (size_t)tri * 3 + 1), cld(m.enodes, (size_t)tri * 3 + 2));
                } else {
                    C gx, gy;
                    node_mean_grad<C>(m, ldA, node, nz, gx, gy);
                    v = gx;
                }
            }
            put(nz, side, v);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// FCT bounds and limiter ratios (lib/oce_adv_tra_fct.F90:96-374).  The cluster of node n at
// level nz is the set of nodes of the elements around n that are wet at nz; with complete
// nod_in_elem2D that is n plus every neighbour m whose edge (n, m) reaches level nz, and an
// element dry at nz contributes -1e3 / +1e3 exactly when nz > nlevels_nod2D_min(n) - 1.
// The column is swept over cluster levels j: after the cluster of level j is formed, layer j - 1
// (which needs the clusters j - 2, j - 1, j) is finished; the surface and bottom layers use
// their own cluster only (the bottom layer L is finished at j = L + 1, where no cluster is
// formed).  TILED stages max(LO, T) and min(LO, T) of the tile and its halo.
#ifndef FESOMTRA_BOUNDS_PF
#define FESOMTRA_BOUNDS_PF 1
#endif
template <bool TILED, int KB>
constexpr bool kBoundsPfFor = FESOMTRA_BOUNDS_PF != 0 && !TILED && KB == 64;
template <bool TILED, int KB, typename S, typename M, typename C, bool GEN = false>
__global__ __launch_bounds__(256, (kWaves<TILED, KB>)) void k_bounds(const MeshCol<M> m, const StepPtrs<S, C> p, double dt_step, int lc) {
    using D = double;
    __shared__ D shi[TILED ? 2 : 1][TILED ? kHtSlots : 1], slo[TILED ? 2 : 1][TILED ? kHtSlots : 1];
    HaloTile<KB, M> ht;
    int n, lane, L, top, nz0, nz1, jb, je, js;  // sweep j = jb..je; levels up to js are staged
    bool act = true;
    if constexpr (TILED) {
        if (!ht.init(m, lc)) return;
        n = ht.n0 + (int)threadIdx.x / KB;
        lane = (int)threadIdx.x % KB;
        act = n < m.N && lane_on(p.mask, lane);
        L = act ? cld(m.nlev_n, n) - 1 : 0;
        top = act ? node_top<GEN>(m, n) : 1;
        nz0 = max(ht.nz_b, top);
        nz1 = min(ht.nz_e, L);
        act = act && nz0 <= nz1;
        jb = max(1, ht.nz_b - 1);
        je = ht.nz_e + 1;
        js = min(ht.nz_e + 1, ht.lmax);
    } else {
        column_lane<KB>(n, lane);
        if (n >= m.N || !lane_on(p.mask, lane)) return;
        L = cld(m.nlev_n, n) - 1;
        top = node_top<GEN>(m, n);
        if (!chunk_layers(L, lc, nz0, nz1)) return;
        nz0 = max(nz0, top);
        if (nz0 > nz1) return;
        jb = max(top, nz0 - 1);
        je = nz1 + 1;
        js = 0;
    }
    const size_t Np = m.Np, Ep = m.Ep;
    const int dry_from = act ? cld(m.nlev_min, n) : 0;  // levels >= this have a dry element around n
    const double dt = node_dt<GEN>(m, act ? n : 0, dt_step);
    NodeEdges<TILED, kDegCacheFor<KB>, GEN> ne;
    if (act) load_edges(m, n, ne);
    else no_edges(ne);
    const LaneLd<KB, D, S> ldT{p.T, lane};
    // LO of a cell with tracer value t (stored as LO in FP64, as LO - T in single precision)
    auto lo_at = [&](size_t cell, D t) -> D {
        if constexpr (std::is_same<C, double>::value) return p.LO[cell * KB + lane];
        else return t + (D)p.LO[cell * KB + lane];
    };
    // max / min of the cluster: the reference's (a > b ? a : b) in FP64; single precision uses the
    // hardware max / min (one instruction; equal but for the choice between signed zeros)
    auto bmax = [](D a, D b) { return std::is_same<C, double>::value ? cmax(a, b) : fmax(a, b); };
    auto bmin = [](D a, D b) { return std::is_same<C, double>::value ? cmin(a, b) : fmin(a, b); };
    // Elements around n whose top lies below level nz contribute the values the reference finds in
    // its shared work array there (m.cq_*, p.cqv, written by k_cavity_aux before this kernel).
    auto cavity_bounds = [&](int nz, D &vmax, D &vmin) {
        if constexpr (!GEN) return;
        if (!m.cq_ptr) return;
        const int qe = cld(m.cq_ptr, n + 1);
        for (int i = cld(m.cq_ptr, n); i < qe; ++i) {
            const int sl = cld(m.cq_ent, i);
            if (nz >= cld(m.cq_top, sl)) continue;
            const size_t r = ((size_t)cld(m.cq_off, sl) + (size_t)(nz - 1)) * 2;
            vmax = bmax(vmax, (D)p.cqv[r * KB + lane]);
            vmin = bmin(vmin, (D)p.cqv[(r + 1) * KB + lane]);
        }
    };
    auto cluster = [&](int nz, int buf, D &vmax, D &vmin, D &own_lo) {
        if constexpr (TILED) {  // own_lo: taken from the staging fetch by the caller
            vmax = shi[buf][threadIdx.x];
            vmin = slo[buf][threadIdx.x];
            for_edges(m, ne, nz, [&](int, bool, int oth) {
                vmax = bmax(vmax, shi[buf][oth * KB + lane]);
                vmin = bmin(vmin, slo[buf][oth * KB + lane]);
            });
        } else {
            const size_t o = (size_t)(nz - 1) * Np;
            const D t0 = ldT(o + n), l0 = lo_at(o + n, t0);
            own_lo = l0;
            vmax = bmax(l0, t0);
            vmin = bmin(l0, t0);
            for_edges_ld<true>(
                m, ne, nz,
                [&](int, bool, int oth) {
                    const D tt = ldT(o + oth);
                    return make_double2(tt, lo_at(o + oth, tt));
                },
                [&](double2 v, int, bool, int) {
                    vmax = bmax(vmax, bmax(v.y, v.x));
                    vmin = bmin(vmin, bmin(v.y, v.x));
                });
        }
        if (nz >= dry_from) {
            vmax = bmax(vmax, D(-1.0e3));
            vmin = bmin(vmin, D(1.0e3));
        }
        cavity_bounds(nz, vmax, vmin);
    };
    // 64-wide untiled form: the next level's cluster values (raw storage values of the node and
    // of its cached neighbours) are loaded while a level is finished, in two register sets
    constexpr bool CPF = kBoundsPfFor<TILED, KB>;
    struct TL {
        S t;
        C l;
    };
    struct CPre {
        TL own;
        TL v[decltype(ne)::kN];
    };
    auto tl_load = [&](size_t o, int node) {
        const size_t q = (o + node) * KB + lane;
        return TL{p.T[q], p.LO[q]};
    };
    auto cl_load = [&](int nz, CPre &c) {
        const size_t o = (size_t)(nz - 1) * Np;
        c.own = tl_load(o, n);
        edges_pre(ne, nz, [&](int, bool, int oth) { return tl_load(o, oth); }, c.v);
    };
    auto cluster_pf = [&](int nz, const CPre &c, D &vmax, D &vmin, D &own_lo) {
        const size_t o = (size_t)(nz - 1) * Np;
        auto lo_of = [](D t, C l) -> D {
            if constexpr (std::is_same<C, double>::value) return (D)l;
            else return t + (D)l;
        };
        const D t0 = ldc<D>(&c.own.t, 0), l0 = lo_of(t0, c.own.l);
        own_lo = l0;
        vmax = bmax(l0, t0);
        vmin = bmin(l0, t0);
        edges_use(m, ne, nz, c.v, [&](int, bool, int oth) { return tl_load(o, oth); },
                  [&](const TL &v, int, bool, int) {
                      const D tt = ldc<D>(&v.t, 0), ll = lo_of(tt, v.l);
                      vmax = bmax(vmax, bmax(ll, tt));
                      vmin = bmin(vmin, bmin(ll, tt));
                  });
        if (nz >= dry_from) {
            vmax = bmax(vmax, D(-1.0e3));
            vmin = bmin(vmin, D(1.0e3));
        }
        cavity_bounds(nz, vmax, vmin);
    };
    // P+/P-, Q+/Q- and the ratios of layer nz with cluster bounds (hi, lw) and low-order
    // value lo; the vertical flux at the layer's bottom is kept for the next layer (av_bot)
    C av_bot = 0;
    // the memory inputs of finish (the cached edges' antidiffusive fluxes, the vertical fluxes at
    // the layer's top and bottom, the volume), loaded before the cluster is formed so that both
    // sets of loads are in flight together
    struct Pre {
        double V;
        C av0, av1;
        C f[decltype(ne)::kN];
    };
    auto preload = [&](int nz, Pre &pre, int node) {
        const size_t cell = (size_t)(nz - 1) * Np + node;
        pre.V = (double)cld(m.areasvol, cell);
        const C a0 = p.adfv[cell * KB + lane];  // read unconditionally: no load inside a branch
        pre.av0 = nz == nz0 ? a0 : av_bot;
        pre.av1 = p.adfv[(cell + Np) * KB + lane];
        const size_t eo = (size_t)(nz - 1) * Ep;
        edges_pre_all(ne, [&](int e, bool, int) { return p.adfh[(eo + e) * KB + lane]; }, pre.f);
    };
    auto finish = [&](int nz, D hi, D lw, D lo, const Pre &pre) {
        const size_t cell = (size_t)(nz - 1) * Np + n;
        const double V = pre.V;
        const D qmax = hi - lo, qmin = lw - lo;
        const C av0 = pre.av0, av1 = pre.av1;
        av_bot = av1;
        double pp = 0.0 + ((double)cmax(C(0), av0) + (double)cmax(C(0), -av1));
        double pm = 0.0 + ((double)cmin(C(0), av0) + (double)cmin(C(0), -av1));
        const size_t eo = (size_t)(nz - 1) * Ep;
        edges_use(
            m, ne, nz, pre.f, [&](int e, bool, int) { return p.adfh[(eo + e) * KB + lane]; },
            [&](C a, int, bool second, int) {
                if (second) a = -a;
                pp = pp + (double)cmax(C(0), a);
                pm = pm + (double)cmin(C(0), a);
            });
        const double fp = pp * dt / V + 1e-16;
        const double fm = pm * dt / V - 1e-16;
        const double rp = (double)qmax / fp, rm = (double)qmin / fm;
        p.Rp[cell * KB + lane] = (C)(rp < 1.0 ? rp : 1.0);
        p.Rm[cell * KB + lane] = (C)(rm < 1.0 ? rm : 1.0);
    };
    // staging registers: the raw values and LO of the fetched level (combined into the cluster
    // values when staged, so that the fetch does not wait for its own loads)
    S t_r[2] = {S(0), S(0)};
    C l_r[2] = {C(0), C(0)};
    auto fetch = [&](int j) {
        ht.each([&](int i, int node) {
            const size_t c = ((size_t)(j - 1) * Np + node) * KB + lane;
            t_r[i] = p.T[c];
            l_r[i] = p.LO[c];
        });
    };
    if constexpr (TILED) fetch(jb);
    D mx2 = 0, mn2 = 0, mx1 = 0, mn1 = 0, mx0 = 0, mn0 = 0;  // clusters of levels j-2, j-1, j
    D lo1 = 0, lo0 = 0;                                      // this column's LO at j-1, j
    // level j of the sweep with the prefetched cluster values cp (CPF; the other forms run the
    // same sequence inline below)
    auto step = [&](int j, const CPre *cp) {
        if constexpr (KB == 64) scalar_fence(ne);
        mx2 = mx1;
        mn2 = mn1;
        mx1 = mx0;
        mn1 = mn0;
        lo1 = lo0;
        const int nz = j - 1;
        const bool fin = nz >= nz0 && nz <= nz1;
        Pre pre;
        if (fin) preload(nz, pre, n);
        if (j <= L && j >= top) cluster_pf(j, *cp, mx0, mn0, lo0);
        if (fin) {
            const bool own = nz == top || nz == L;
            finish(nz, own ? mx1 : bmax(bmax(mx2, mx1), mx0), own ? mn1 : bmin(bmin(mn2, mn1), mn0), lo1, pre);
        }
    };
    if constexpr (CPF) {
        // two register sets, alternating: the cluster values of level j + 1 load during level j
        CPre c0, c1;
        if (jb <= L) cl_load(jb, c0);
        for (int j = jb; j <= je; j += 2) {
            if (j + 1 <= L) cl_load(j + 1, c1);
            step(j, &c0);
            if (j + 1 <= je) {
                if (j + 2 <= L) cl_load(j + 2, c0);
                step(j + 1, &c1);
            }
        }
        return;
    }
    for (int j = jb; j <= je; ++j) {
        const int buf = j & 1;
        const int nz = j - 1;
        const bool fin = act && nz >= nz0 && nz <= nz1;
        Pre pre;
        D lo_j = 0;
        if constexpr (TILED) {
            // the layer's own inputs are requested before the next level's fetch: loads complete
            // in issue order, so requested after it they would make the layer wait for the fetch.
            // Requested by every thread (a column with nothing to finish reads a cell in bounds
            // and ignores it): a load inside a branch would make the staging below wait for it.
            preload(fin ? nz : 1, pre, fin ? n : ht.snode[0]);
            if (j <= js) {
                ht.each([&](int i, int) {
                    const D t = ldc<D>(&t_r[i], 0);
                    const D l = std::is_same<C, double>::value ? (D)l_r[i] : t + (D)l_r[i];
                    shi[buf][threadIdx.x + 256 * i] = bmax(l, t);
                    slo[buf][threadIdx.x + 256 * i] = bmin(l, t);
                    if (i == 0) lo_j = l;
                });
                __syncthreads();
                if (j < js) fetch(j + 1);
            }
            if (!act) continue;
        }
        mx2 = mx1;
        mn2 = mn1;
        mx1 = mx0;
        mn1 = mn0;
        lo1 = lo0;
        if constexpr (!TILED)
            if (fin) preload(nz, pre, n);
        if (j <= L && j >= top) {
            cluster(j, buf, mx0, mn0, lo0);
            if constexpr (TILED) lo0 = lo_j;
        }
        if (fin) {
            const bool own = nz == top || nz == L;
            finish(nz, own ? mx1 : bmax(bmax(mx2, mx1), mx0), own ? mn1 : bmin(bmin(mn2, mn1), mn0), lo1, pre);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Limiting (b3) + flux divergence (oce_tra_adv_flux2dtracer) + caller update.  With acc_h /
// acc_v the tendencies are accumulated there (do_oce_adv_tra contract); otherwise values +=
// (dttf_h + dttf_v) / hnode_new in place.  dbg_h / dbg_v receive the limited fluxes if set.
// acc_* and dbg_* hold one tracer: they are only passed with a single-lane mask.
// Per-edge inputs of the flux divergence: antidiffusive flux, the far node's limiter ratios,
// and (single precision without limiter) the far node's valuesAB and the transport.
template <typename C>
struct EdgeVals {
    C f, rp, rm;
    double ao, q;
};
#ifndef FESOMTRA_FINAL_PRE
#define FESOMTRA_FINAL_PRE 1
#endif
// Untiled update kernel: the edge inputs of a layer are loaded before the column's own inputs
// are used (one memory round per layer instead of two), where the registers allow it (one column
// per wavefront, or single-precision values; narrow FP64 columns would spill).
constexpr bool kFinalPre = FESOMTRA_FINAL_PRE != 0;
template <bool TILED, int KB, typename C>
constexpr bool kFinalPreFor = kFinalPre && !TILED && (KB == 64 || !std::is_same<C, double>::value);
// AB2: also writes the next step's valuesAB and valuesold (a separate instantiation, so that the
// other runs keep their register budget).  LEAN (one column per wavefront): a pass without the
// drop-in's accumulators or kept fluxes, compiled without them, whose pointer tests would
// otherwise hold scalar registers through the level loop.
template <bool TILED, bool FCT, int KB, typename S, typename M, typename C, bool AB2 = false, bool GEN = false,
          bool LEAN = false>
__global__ __launch_bounds__(256, (kWaves<TILED, KB>)) void k_final(const MeshCol<M> m, const StepPtrs<S, C> p, double dt_step, int lc) {
    static_assert(!TILED || FCT, "tiles stage the limiter ratios");
    constexpr bool PRE = kFinalPreFor<TILED, KB, C>;
    __shared__ C sRp[TILED ? 2 : 1][TILED ? kHtSlots : 1], sRm[TILED ? 2 : 1][TILED ? kHtSlots : 1];
    HaloTile<KB, M> ht;
    int n, lane, L, ctop, nz0, nz1, lv0, lv1;  // ctop: the column's top layer
    bool act = true;
    if constexpr (TILED) {
        if (!ht.init(m, lc)) return;
        n = ht.n0 + (int)threadIdx.x / KB;
        lane = (int)threadIdx.x % KB;
        act = n < m.N && lane_on(p.mask, lane);
        L = act ? cld(m.nlev_n, n) - 1 : 0;
        ctop = act ? node_top<GEN>(m, n) : 1;
        nz0 = max(ht.nz_b, ctop);
        nz1 = min(ht.nz_e, L);
        act = act && nz0 <= nz1;
        lv0 = ht.nz_b;
        lv1 = ht.nz_e;
    } else {
        column_lane<KB>(n, lane);
        if (n >= m.N || !lane_on(p.mask, lane)) return;
        L = cld(m.nlev_n, n) - 1;
        ctop = node_top<GEN>(m, n);
        if (!chunk_layers(L, lc, nz0, nz1)) return;
        nz0 = max(nz0, ctop);
        if (nz0 > nz1) return;
        lv0 = nz0;
        lv1 = nz1;
    }
    const double dt = node_dt<GEN>(m, act ? n : 0, dt_step);
    const size_t Np = m.Np, Ep = m.Ep;
    NodeEdges<TILED, kDegCacheFor<KB>, GEN> ne;
    if (act) load_edges(m, n, ne);
    else no_edges(ne);
    // limited flux at interface j given R at layers j-1 (rp0, rm0) and j (rp1, rm1)
    auto limit = [&](C a, int j, C rp0, C rm0, C rp1, C rm1) -> C {
        C ae = 1;
        if (j == ctop) ae = a >= C(0) ? cmin(ae, rp1) : cmin(ae, rm1);
        else if (a >= C(0)) {
            ae = cmin(ae, rm0);
            ae = cmin(ae, rp1);
        } else {
            ae = cmin(ae, rp0);
            ae = cmin(ae, rm1);
        }
        return ae * a;
    };
    auto X = [&](size_t cell) { return cell * KB + lane; };
    C top = 0, rp0 = 0, rm0 = 0, rp1 = 0, rm1 = 0;
    const size_t c0 = (size_t)(nz0 - 1) * Np + n;
    if (act) {
        const C a0 = p.adfv[X(c0)];
        if (FCT) {
            rp1 = p.Rp[X(c0)];
            rm1 = p.Rm[X(c0)];
            const C rpa = nz0 > ctop ? p.Rp[X(c0 - Np)] : C(0);
            const C rma = nz0 > ctop ? p.Rm[X(c0 - Np)] : C(0);
            top = limit(a0, nz0, rpa, rma, rp1, rm1);
            if (!LEAN && p.dbg_v) p.dbg_v[c0] = top;
        } else {
            top = a0;
        }
    }
    C sp[2], sm[2];
    if constexpr (TILED) {
        ht.each([&](int i, int node) {
            const size_t c = (size_t)(lv0 - 1) * Np + node;
            sp[i] = p.Rp[X(c)];
            sm[i] = p.Rm[X(c)];
        });
    }
    // Tiled form: a layer's own inputs are requested before the next layer's staging fetch (its
    // cached edges' fluxes and its R one layer down even before the staging barrier).  Loads
    // complete in issue order, so requested after that fetch they would make the layer wait for
    // the fetch too; requested before, the fetch stays in flight while the layer is computed.
    struct OwnIn {
        double V, hn, h, tval, lo;
        C ab, rpd, rmd;
        C f[decltype(ne)::kN];
    };
    OwnIn own;
    auto own_edges = [&](int nz, int node) {
        const size_t cell = (size_t)(nz - 1) * Np + node, eo = (size_t)(nz - 1) * Ep;
        if constexpr (FCT) {
            const size_t cd = nz + 1 <= L ? cell + Np : cell;  // R one layer down (unused at the bottom)
            own.rpd = p.Rp[X(cd)];
            own.rmd = p.Rm[X(cd)];
        }
        edges_pre_all(ne, [&](int e, bool, int) { return p.adfh[X(eo + e)]; }, own.f);
    };
    auto own_col = [&](int nz, int node) {
        const size_t cell = (size_t)(nz - 1) * Np + node;
        own.V = (double)cld(m.areasvol, cell);
        own.hn = (double)cld(m.hnode_new, cell);
        own.h = FCT ? (double)cld(m.hnode, cell) : 0.0;
        own.tval = (double)ldc<C>(p.T, X(cell));
        own.ab = p.adfv[X(cell + Np)];
        if constexpr (FCT) {
            // the stored LO (single precision: the increment and its remainder, combined at use)
            own.lo = std::is_same<C, double>::value
                         ? (double)p.LO[X(cell)]
                         : (double)p.LO[X(cell)] + (p.LOe ? (double)p.LOe[X(cell)] : 0.0);
        }
    };
    for (int nz = lv0; nz <= lv1; ++nz) {
        const int buf = nz & 1;
        if constexpr (KB == 64) scalar_fence(ne);
        if constexpr (TILED) {
            // requested by every thread (one with nothing to update reads a cell in bounds and
            // ignores it): a load inside a branch would make the staging below wait for it
            const bool live = act && nz <= nz1 && (!GEN || nz >= nz0);
            const int onz = live ? nz : 1, onode = live ? n : ht.snode[0];
            own_edges(onz, onode);
            ht.each([&](int i, int) {
                sRp[buf][threadIdx.x + 256 * i] = sp[i];
                sRm[buf][threadIdx.x + 256 * i] = sm[i];
            });
            __syncthreads();
            own_col(onz, onode);
            if (nz < lv1) {
                ht.each([&](int i, int node) {
                    const size_t c = (size_t)nz * Np + node;
                    sp[i] = p.Rp[X(c)];
                    sm[i] = p.Rm[X(c)];
                });
            }
            if (!live) continue;
        }
        const size_t cell = (size_t)(nz - 1) * Np + n;
        const size_t eo = (size_t)(nz - 1) * Ep, no = (size_t)(nz - 1) * Np;
        // single precision without limiter: the stored flux is high-order minus upwind; the
        // upwind part of valuesAB is added here in double
        constexpr bool kAddUpw = !FCT && !std::is_same<C, double>::value;
        // untiled: the cached edges' inputs are loaded first, in flight with the column's own
        auto edge_load = [&](int e, bool, int oth) {
            EdgeVals<C> v{p.adfh[X(eo + e)], C(0), C(0), 0.0, 0.0};
            if (FCT) {
                v.rp = p.Rp[X(no + oth)];
                v.rm = p.Rm[X(no + oth)];
            }
            if (kAddUpw) {
                v.ao = ldc<double>(p.TAB, X(no + oth));
                v.q = cld(m.q, lane_member<GEN>(m, p, lane).at(eo + e));
            }
            return v;
        };
        EdgeVals<C> pre[PRE ? decltype(ne)::kN : 1];
        if constexpr (PRE) edges_pre(ne, nz, edge_load, pre);
        const double V = TILED ? own.V : (double)cld(m.areasvol, cell);
        const double hn = TILED ? own.hn : (double)cld(m.hnode_new, cell);
        const double h = TILED ? own.h : FCT ? (double)cld(m.hnode, cell) : 0.0;
        const double tval = TILED ? own.tval : (double)ldc<C>(p.T, X(cell));
        const C ab = TILED ? own.ab : p.adfv[X(cell + Np)];
        C bot;
        double dv = (!LEAN && p.acc_v && !p.acc_init) ? p.acc_v[cell] : 0.0;
        if (FCT) {
            // single precision: LO = T + increment (+ the increment's remainder with compensated
            // storage, so that the low-order change of the content is the fluxes' to FP64 round-off)
            const double lo = std::is_same<C, double>::value
                                  ? (TILED ? own.lo : (double)p.LO[X(cell)])
                                  : tval + (TILED ? own.lo
                                                  : ((double)p.LO[X(cell)] + (p.LOe ? (double)p.LOe[X(cell)] : 0.0)));
            dv = dv - tval * h + lo * hn;
            rp0 = rp1;
            rm0 = rm1;
            if (nz + 1 <= L) {
                if (TILED) {
                    rp1 = own.rpd;
                    rm1 = own.rmd;
                } else {
                    rp1 = p.Rp[X(cell + Np)];
                    rm1 = p.Rm[X(cell + Np)];
                }
                bot = limit(ab, nz + 1, rp0, rm0, rp1, rm1);
            } else {
                bot = ab;  // bottom interface: never limited (always zero)
            }
            if (!LEAN && p.dbg_v) p.dbg_v[cell + Np] = bot;
        } else {
            bot = ab;
        }
        const double yV = kDivReuse<KB> ? 1.0 / V : 0.0;
        dv = dv + (kDivReuse<KB> ? div_reuse(((double)top - (double)bot) * dt, V, yV) : ((double)top - (double)bot) * dt / V);
        double dh = (!LEAN && p.acc_h && !p.acc_init) ? p.acc_h[cell] : 0.0;
        const C rpn = FCT ? rp0 : C(0), rmn = FCT ? rm0 : C(0);
        const double an = kAddUpw ? ldc<double>(p.TAB, X(cell)) : 0.0;
        auto add_edge = [&](const EdgeVals<C> &v, int e, bool second) {
            C f = v.f;
            if (FCT) {
                // R of both end nodes (lib/oce_adv_tra_fct.F90:455-468)
                const C rp_1 = second ? v.rp : rpn, rm_1 = second ? v.rm : rmn;  // edges(1)
                const C rp_2 = second ? rpn : v.rp, rm_2 = second ? rmn : v.rm;  // edges(2)
                C ae = 1;
                if (f >= C(0)) {
                    ae = cmin(ae, rp_1);
                    ae = cmin(ae, rm_2);
                } else {
                    ae = cmin(ae, rm_1);
                    ae = cmin(ae, rp_2);
                }
                f = ae * f;
                if (!LEAN && p.dbg_h && !second) p.dbg_h[eo + e] = f;
            }
            double fd = (double)f;
            if (kAddUpw) fd = fd + upw1_edge<double>(second ? v.ao : an, second ? an : v.ao, v.q);
            const double x = kDivReuse<KB> ? div_reuse(fd * dt, V, yV) : fd * dt / V;
            dh = second ? dh - x : dh + x;
        };
        if constexpr (TILED) {
            // antidiffusive fluxes requested before the barrier; limiter ratios from the tile
            edges_use(m, ne, nz, own.f, [&](int e, bool, int) { return p.adfh[X(eo + e)]; },
                      [&](C f, int e, bool second, int oth) {
                          add_edge(EdgeVals<C>{f, sRp[buf][oth * KB + lane], sRm[buf][oth * KB + lane], 0.0, 0.0}, e,
                                   second);
                      });
        } else if constexpr (PRE) {
            edges_use(m, ne, nz, pre, edge_load, [&](const EdgeVals<C> &v, int e, bool second, int) { add_edge(v, e, second); });
        } else {
            for_edges_ld<kTwoPhase<KB>>(m, ne, nz, edge_load,
                                        [&](const EdgeVals<C> &v, int e, bool second, int) { add_edge(v, e, second); });
        }
        if (!LEAN && p.acc_h) {
            p.acc_h[cell] = dh;
            p.acc_v[cell] = dv;
        } else if (p.resid) {
            // compensated storage: the value is stored + remainder; the step's increment is added
            // to both (in double), the sum rounded to storage and the new remainder kept, so the
            // content sum(T h V) changes by the fluxes only (no storage rounding accumulates)
            const double full = tval + (double)p.resid[X(cell)] + (dh + dv) / hn;
            std_(p.T, X(cell), full);
            p.resid[X(cell)] = (float)(full - ldc<double>(p.T, X(cell)));
        } else if (p.seed != 0u && !std::is_same<S, double>::value) {
            const uint32_t step = p.seed_dev ? *p.seed_dev : p.seed;
            st_stochastic(p.T, X(cell), tval + (dh + dv) / hn,
                          unit_hash((uint32_t)(p.tid0 + lane), (uint32_t)cell, step));
        } else {
            std_(p.T, X(cell), tval + (dh + dv) / hn);
        }
        if constexpr (AB2) {
            if (!p.acc_h) {
                // FESOM2's AB2 interpolation for the next step from the value just stored and this
                // step's (lib/oce_modules.F90:75): valuesAB = -(0.5+eps) valuesold + (1.5+eps)
                // values with valuesold := the values this step started from
                std_(p.ab_next, X(cell), -p.ab_c0 * tval + p.ab_c1 * ldc<double>(p.T, X(cell)));
                std_(p.ab_old, X(cell), tval);
            }
        }
        top = bot;
    }
}

// ---------------------------------------------------------------------------------------------
// Fused limiter: bounds, ratios, limiting, flux divergence and update in one kernel
// (lib/oce_adv_tra_fct.F90:96-471, lib/oce_adv_tra_driver.F90:298-450, src/fesom_analytic.F90:419-437)
// for tracer blocks of width KB <= 4, FP64, meshes without cavities, per-node time steps or
// ensembles, one level chunk per column.
//
// The update of a tile node needs the limiter ratios R of its one-ring neighbours, which need
// the cluster values max(LO, T) / min(LO, T) of their own neighbours: the tile's second ring.  A
// workgroup owns a tile of 128 / KB consecutive nodes; its 256 / KB node threads hold the tile's
// nodes and then its first ring, so every thread forms the ratios of one column (the first-ring
// columns a second time, as the neighbouring tiles do too) and only the tile's threads go on to
// limit and update.  Each thread stages two slots: its own column and one second-ring node.  R
// never goes through memory and the antidiffusive fluxes, the vertical fluxes, LO and the volumes
// are read once for both phases.  Sweep step j: stage the cluster values of level j and the
// ratios of layer j - 2; one barrier; fetch level j + 1; form the ratios of layer j - 1 (clusters
// j - 2 .. j, as k_bounds); limit and update layer j - 2 (R of layers j - 3 .. j - 1 of the
// column, R of layer j - 2 at the neighbours, as k_final).  Every expression and summation order
// is that of k_bounds and k_final, so results are bit-identical.  The updated values go to a
// second buffer (neighbouring tiles read the old values of their rings while a tile updates).
#ifndef FESOMTRA_MIN_WAVES_LIMIT
#define FESOMTRA_MIN_WAVES_LIMIT 3
#endif
template <int KB, typename M>
struct SplitTile {
    static constexpr int NT = 256 / KB, TCS = 128 / KB;
    int t, n0, lmax, lstage, hb, nr1;
    int node;  // this thread's column: a tile node (tile), a first-ring node, or none (-1)
    bool tile;
    int r2node;  // the second-ring node staged in this thread's second slot
    bool r2on;
    __device__ __forceinline__ bool init(const MeshCol<M> &m, int col) {
        t = xcd_block();
        if (t >= m.rt_n) return false;
        n0 = t * TCS;
        lmax = cld(m.rt_lmax, t);
        lstage = min(lmax + 1, cld(m.rt_smax, t));
        hb = cld(m.rt_hptr, t);
        const int H = cld(m.rt_hptr, t + 1) - hb;
        nr1 = cld(m.rt_nr1, t);
        tile = col < TCS;
        if (tile) node = n0 + col < m.N ? n0 + col : -1;
        else node = col - TCS < nr1 ? cld(m.rt_halo, hb + col - TCS) : -1;
        r2on = col < H - nr1;
        r2node = r2on ? cld(m.rt_halo, hb + nr1 + col) : 0;
        return true;
    }
};

template <int KB, typename S, typename M, typename C>
__global__ __launch_bounds__(256, FESOMTRA_MIN_WAVES_LIMIT) void k_limit(const MeshCol<M> m, const StepPtrs<S, C> p, double dt) {
    static_assert(std::is_same<C, double>::value && std::is_same<S, double>::value, "FP64 only");
    using D = double;
    constexpr int NT = SplitTile<KB, M>::NT, NC = kDegCacheFor<KB>;
    __shared__ D shi[2][2 * 256], slo[2][2 * 256];  // cluster values of a level, by level parity
    __shared__ C sRp[2][256], sRm[2][256];           // R of a layer (tile and first ring), by layer parity
    SplitTile<KB, M> st;
    const int col = (int)threadIdx.x / KB, lane = (int)threadIdx.x % KB;
    if (!st.init(m, col)) return;
    const size_t Np = m.Np, Ep = m.Ep;
    auto X = [&](size_t cell) { return cell * KB + lane; };
    const int n = st.node >= 0 ? st.node : 0;  // a column in bounds for threads without one
    bool act = st.node >= 0 && lane_on(p.mask, lane);
    const int Lc = act ? cld(m.nlev_n, n) - 1 : 0;
    const int L = st.tile ? Lc : min(Lc, st.lmax);  // layers whose ratios this thread forms
    act = act && L >= 1;
    const bool upd_col = act && st.tile;
    const int dry_from = act ? cld(m.nlev_min, n) : 0;
    // the column's entries: the first NC in registers (edge * 2 + second, and deepest level |
    // far slot << 16), the rest from memory; edges(nz, f) calls f(entry, edge, second, slot) in
    // reference order
    const int eb = act ? cld(m.ne_ptr, n) : 0;
    const int deg = act ? cld(m.ne_ptr, n + 1) - eb : 0;
    const int eoff = act ? cld(m.rt_eoff, st.t * NT + col) : 0;
    int ce2[NC], ctop[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const bool on = i < deg;
        ce2[i] = on ? cld(m.ne_e2, eb + i) : 0;
        ctop[i] = on ? (cld(m.rt_loc, eoff + i) << 16) | cld(m.ne_top, eb + i) : 0;
    }
    auto edges = [&](int nz, auto &&f) {
#pragma unroll
        for (int i = 0; i < NC; ++i)
            if (nz <= (ctop[i] & 0xffff)) f(i, ce2[i] >> 1, (ce2[i] & 1) != 0, ctop[i] >> 16);
        for (int i = NC; i < deg; ++i) {
            if (nz > cld(m.ne_top, eb + i)) continue;
            const int e2 = cld(m.ne_e2, eb + i);
            f(i, e2 >> 1, (e2 & 1) != 0, cld(m.rt_loc, eoff + i));
        }
    };
    // ratios of layer nz with cluster bounds (hi, lw), low-order value lo, vertical fluxes av0 /
    // av1 and volume V; the column's antidiffusive edge fluxes come through use_edges (k_bounds'
    // finish)
    auto ratios = [&](D hi, D lw, D lo, C av0, C av1, double V, auto &&use_edges, C &rp, C &rm) {
        const D qmax = hi - lo, qmin = lw - lo;
        double pp = 0.0 + ((double)cmax(C(0), av0) + (double)cmax(C(0), -av1));
        double pm = 0.0 + ((double)cmin(C(0), av0) + (double)cmin(C(0), -av1));
        use_edges([&](C a, bool second) {
            if (second) a = -a;
            pp = pp + (double)cmax(C(0), a);
            pm = pm + (double)cmin(C(0), a);
        });
        const double fp = pp * dt / V + 1e-16;
        const double fm = pm * dt / V - 1e-16;
        const double r_p = (double)qmax / fp, r_m = (double)qmin / fm;
        rp = (C)(r_p < 1.0 ? r_p : 1.0);
        rm = (C)(r_m < 1.0 ? r_m : 1.0);
    };
    // limited flux at interface j from R at layers j - 1 (rp0, rm0) and j (rp1, rm1) (k_final)
    auto limit = [&](C a, int j, C rp0, C rm0, C rp1, C rm1) -> C {
        C ae = 1;
        if (j == 1) ae = a >= C(0) ? cmin(ae, rp1) : cmin(ae, rm1);
        else if (a >= C(0)) {
            ae = cmin(ae, rm0);
            ae = cmin(ae, rp1);
        } else {
            ae = cmin(ae, rp0);
            ae = cmin(ae, rm1);
        }
        return ae * a;
    };
    // staging registers of the next level: raw values and LO of both slots (combined when staged,
    // so that the fetch does not wait for its own loads)
    D t_r[2] = {0, 0}, l_r[2] = {0, 0};
    const int snode[2] = {n, st.r2node};
    const bool son[2] = {st.node >= 0, st.r2on};
    auto fetch = [&](int j) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const size_t c = (size_t)(j - 1) * Np + snode[i];
            t_r[i] = p.T[X(c)];
            l_r[i] = p.LO[X(c)];
        }
    };
    // sweep state: clusters of levels j - 2, j - 1, j and LO at j - 1, j; the column's inputs of
    // layer j - 2 kept from its ratio step for the update (antidiffusive and vertical fluxes,
    // volume, LO)
    D mx2 = 0, mn2 = 0, mx1 = 0, mn1 = 0, mx0 = 0, mn0 = 0, lo1 = 0, lo0 = 0;
    C rpB = 0, rmB = 0;  // R of layer j - 2 of the column (formed in the previous step)
    C f_k[NC], av0_k = 0, av1_k = 0, av_bot = 0, topf = 0;
    double V_k = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) f_k[i] = 0;
    D lo_k = 0;
    if (st.lstage >= 1) fetch(1);
    for (int j = 1; j <= st.lmax + 2; ++j) {
        const int buf = j & 1, rbuf = j & 1;  // R of layer j - 2 goes to parity (j - 2) & 1 = j & 1
        // this step's own inputs, requested before the next level's fetch (loads complete in
        // issue order) and by every thread (a thread with nothing to do reads cells in bounds and
        // ignores them, so no load sits inside a branch): the ratio step's (layer j - 1) and the
        // update's (layer j - 2)
        const int nz = j - 1, nf = j - 2;
        const bool fin = act && nz >= 1 && nz <= L;
        const bool upd = upd_col && nf >= 1 && nf <= L;
        C f_n[NC], av0_n, av1_n;
        double V_n;
        {
            const int z = fin ? nz : 1;
            const size_t cell = (size_t)(z - 1) * Np + n, eo = (size_t)(z - 1) * Ep;
            V_n = (double)cld(m.areasvol, cell);
            const C a0 = p.adfv[X(cell)];
            av0_n = z == 1 ? a0 : av_bot;
            av1_n = p.adfv[X(cell + Np)];
#pragma unroll
            for (int i = 0; i < NC; ++i) f_n[i] = p.adfh[X(eo + (ce2[i] >> 1))];
        }
        double u_hn, u_hh, u_t, u_dv = 0.0, u_dh = 0.0;
        {
            const size_t cell = (size_t)((upd ? nf : 1) - 1) * Np + n;
            u_hn = (double)cld(m.hnode_new, cell);
            u_hh = (double)cld(m.hnode, cell);
            u_t = (double)ldc<C>(p.T, X(cell));
            if (p.acc_h && !p.acc_init) {
                u_dv = p.acc_v[cell];
                u_dh = p.acc_h[cell];
            }
        }
        D lo_j = 0;
        if (j <= st.lstage) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                if (!son[i]) continue;
                const D t = t_r[i], l = l_r[i];
                shi[buf][threadIdx.x + 256 * i] = cmax(l, t);
                slo[buf][threadIdx.x + 256 * i] = cmin(l, t);
            }
            lo_j = l_r[0];
        }
        if (j >= 3) {
            // R of layer j - 2 of this thread's column (slot col)
            sRp[rbuf][threadIdx.x] = rpB;
            sRm[rbuf][threadIdx.x] = rmB;
        }
        __syncthreads();
        if (j < st.lstage) fetch(j + 1);
        // ---- ratios of layer j - 1 ----
        mx2 = mx1; mn2 = mn1; mx1 = mx0; mn1 = mn0; lo1 = lo0;
        C rpN = 0, rmN = 0;
        if (act && j <= Lc && j <= st.lstage) {
            mx0 = shi[buf][threadIdx.x];
            mn0 = slo[buf][threadIdx.x];
            edges(j, [&](int, int, bool, int sl) {
                mx0 = cmax(mx0, shi[buf][sl * KB + lane]);
                mn0 = cmin(mn0, slo[buf][sl * KB + lane]);
            });
            if (j >= dry_from) {
                mx0 = cmax(mx0, D(-1.0e3));
                mn0 = cmin(mn0, D(1.0e3));
            }
            lo0 = lo_j;
        }
        if (fin) {
            const bool own = nz == 1 || nz == Lc;
            const size_t eo = (size_t)(nz - 1) * Ep;
            ratios(own ? mx1 : cmax(cmax(mx2, mx1), mx0), own ? mn1 : cmin(cmin(mn2, mn1), mn0), lo1, av0_n, av1_n,
                   V_n,
                   [&](auto &&g) {
                       edges(nz, [&](int i, int e, bool second, int) {
                           g(i < NC ? f_n[i < NC ? i : 0] : p.adfh[X(eo + e)], second);
                       });
                   },
                   rpN, rmN);
            av_bot = av1_n;
        }
        // ---- limiting and update of layer j - 2 ----
        if (upd) {
            const size_t cell = (size_t)(nf - 1) * Np + n;
            const size_t eo = (size_t)(nf - 1) * Ep;
            const int fb = nf & 1;
            // R of this column: layer nf (rpB), nf + 1 (rpN, formed above); the top interface's
            // limited flux was formed with R of nf - 1 by the previous update (topf)
            if (nf == 1) {
                topf = limit(av0_k, 1, C(0), C(0), rpB, rmB);
                if (p.dbg_v) p.dbg_v[cell] = topf;
            }
            C bot;
            if (nf + 1 <= L) {
                bot = limit(av1_k, nf + 1, rpB, rmB, rpN, rmN);
            } else {
                bot = av1_k;  // bottom interface: never limited (always zero)
            }
            if (p.dbg_v) p.dbg_v[cell + Np] = bot;
            const double V = V_k;
            const double hn = u_hn, hh = u_hh, tval = u_t;
            double dv = u_dv;
            dv = dv - tval * hh + (double)lo_k * hn;
            dv = dv + ((double)topf - (double)bot) * dt / V;
            double dh = u_dh;
            const C rpn = rpB, rmn = rmB;
            edges(nf, [&](int i, int e, bool second, int oth) {
                C f = i < NC ? f_k[i < NC ? i : 0] : p.adfh[X(eo + e)];
                const C rpo = sRp[fb][oth * KB + lane], rmo = sRm[fb][oth * KB + lane];
                const C rp_1 = second ? rpo : rpn, rm_1 = second ? rmo : rmn;  // edges(1)
                const C rp_2 = second ? rpn : rpo, rm_2 = second ? rmn : rmo;  // edges(2)
                C ae = 1;
                if (f >= C(0)) {
                    ae = cmin(ae, rp_1);
                    ae = cmin(ae, rm_2);
                } else {
                    ae = cmin(ae, rm_1);
                    ae = cmin(ae, rp_2);
                }
                f = ae * f;
                if (p.dbg_h && !second) p.dbg_h[eo + e] = f;
                const double x = (double)f * dt / V;
                dh = second ? dh - x : dh + x;
            });
            if (p.acc_h) {
                p.acc_h[cell] = dh;
                p.acc_v[cell] = dv;
            } else {
                std_(p.Tout, X(cell), tval + (dh + dv) / hn);
            }
            topf = bot;
        }
        // shift: the ratio step's inputs of layer j - 1 become the update's of the next step
        if (fin) {
#pragma unroll
            for (int i = 0; i < NC; ++i) f_k[i] = f_n[i];
            av0_k = av0_n;
            av1_k = av1_n;
            V_k = V_n;
            lo_k = lo1;
        }
        rpB = rpN;
        rmB = rmN;
    }
}

// ---------------------------------------------------------------------------------------------
// Renumbering / transposition between the reference layout (nl-1, nod2D) in reference order
// and lane `lane` of a device tracer block [k][Np][kb] in device order.
// dir 0: ref -> dev, dir 1: dev -> ref.
// The reference layout is node-major (a node's layers contiguous, nodes in reference order), the
// device layout level-major (a level's nodes contiguous, in device order): a transpose with a
// permutation.  A workgroup takes 32 device nodes x up to 64 layers (one chunk covers the columns
// of most meshes) through LDS, so both sides are read and written in contiguous runs (a node's
// layers on the reference side, 32 nodes of a level on the device side) whatever the
// permutation.  The cells of a chunk are numbered densely (no idle threads for a chunk shorter
// than 64 layers).  With a random node numbering one side is 376-byte runs at random places:
// 32-node tiles (17 KB of LDS, 9 workgroups per CU) keep more of them in flight than 64-node tiles
// (0.245 against 0.311 ms per 225 MB tracer of the 1000x1000x48 mesh).
#ifndef FESOMTRA_RELAY_NODES
#define FESOMTRA_RELAY_NODES 32
#endif
constexpr int kRelayNodes = FESOMTRA_RELAY_NODES, kRelayLevels = 64;
template <typename S>
__global__ __launch_bounds__(256) void k_relayout(const double *__restrict__ ref_in, S *__restrict__ dev_out,
                                                  const S *__restrict__ dev_in, double *__restrict__ ref_out,
                                                  const int *__restrict__ perm, int N, int L, size_t Np, int kb,
                                                  int lane, int dir, float *__restrict__ res) {
    __shared__ double tile[kRelayNodes * (kRelayLevels + 1)];
    const int n0 = (int)blockIdx.x * kRelayNodes, k0 = (int)blockIdx.y * kRelayLevels;
    const int nn_n = min(kRelayNodes, N - n0), nk = min(kRelayLevels, L - k0);
    const int cells = nn_n * nk, ld = nk + 1;  // tile row of a node: its nk layers (+1 against bank conflicts)
    // res (compensated storage): the rounding remainder v - stored, in FP32 beside the stored value
    if (dir == 0) {
        for (int i = (int)threadIdx.x; i < cells; i += (int)blockDim.x) {  // node-major reads
            const int nn = i / nk, kk = i - nn * nk;
            tile[nn * ld + kk] = ref_in[(size_t)perm[n0 + nn] * L + k0 + kk];
        }
        __syncthreads();
        for (int i = (int)threadIdx.x; i < cells; i += (int)blockDim.x) {  // level-major writes
            const int kk = i / nn_n, nn = i - kk * nn_n;
            const size_t d = ((size_t)(k0 + kk) * Np + n0 + nn) * kb + lane;
            const double v = tile[nn * ld + kk];
            std_(dev_out, d, v);
            if (res) res[d] = (float)(v - ldc<double>(dev_out, d));
        }
    } else {
        for (int i = (int)threadIdx.x; i < cells; i += (int)blockDim.x) {  // level-major reads
            const int kk = i / nn_n, nn = i - kk * nn_n;
            const size_t d = ((size_t)(k0 + kk) * Np + n0 + nn) * kb + lane;
            tile[nn * ld + kk] = ldc<double>(dev_in, d) + (res ? (double)res[d] : 0.0);
        }
        __syncthreads();
        for (int i = (int)threadIdx.x; i < cells; i += (int)blockDim.x) {  // node-major writes
            const int nn = i / nk, kk = i - nn * nk;
            ref_out[(size_t)perm[n0 + nn] * L + k0 + kk] = tile[nn * ld + kk];
        }
    }
}

// FESOM2's second-order Adams-Bashforth interpolation of a tracer block before a step:
// valuesAB = -(0.5+eps)*valuesold + (1.5+eps)*values, then valuesold := values (eps of
// lib/oce_modules.F90:75, valuesold of lib/MOD_TRACER.F90:13; c0 = 0.5+eps, c1 = 1.5+eps).  In
// double, stored at the storage precision (FP64: the Fortran expression's rounding exactly).
template <typename S>
__global__ __launch_bounds__(256) void k_ab2(const S *__restrict__ T, S *__restrict__ AB, S *__restrict__ old, size_t n,
                                             double c0, double c1) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double t = ldc<double>(T, i), o = ldc<double>(old, i);
        std_(AB, i, -c0 * o + c1 * t);
        old[i] = T[i];
    }
}

// Stats over one tracer (lane of a block): per-block partial min / max / sum / volume sum over
// wet layers (or all layers).
template <typename S, typename M>
__global__ __launch_bounds__(256) void k_stats_partial(const MeshCol<M> m, const S *__restrict__ T, const float *__restrict__ res,
                                                       int kb, int lane, int all_levels, double4 *__restrict__ part) {
    __shared__ double smin[256], smax[256], ssum[256], svol[256];
    double mn = __builtin_huge_val(), mx = -__builtin_huge_val(), sm = 0.0, sv = 0.0;
    const size_t total = (size_t)m.L * m.N;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int k = (int)(i / m.N), n = (int)(i - (size_t)k * m.N);
        const bool wet = k + 1 <= cld(m.nlev_n, n) - 1 && k + 1 >= node_top(m, n);
        if (!all_levels && !wet) continue;
        const size_t ix = (size_t)k * m.Np + n;
        const double v = ldc<double>(T, ix * kb + lane) + (res ? (double)res[ix * kb + lane] : 0.0);
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
        sm += v;
        if (wet) sv += v * (double)cld(m.hnode, ix) * (double)cld(m.areasvol, ix);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    svol[threadIdx.x] = sv;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = smin[threadIdx.x] < smin[threadIdx.x + s] ? smin[threadIdx.x] : smin[threadIdx.x + s];
            smax[threadIdx.x] = smax[threadIdx.x] > smax[threadIdx.x + s] ? smax[threadIdx.x] : smax[threadIdx.x + s];
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            svol[threadIdx.x] += svol[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double4(smin[0], smax[0], ssum[0], svol[0]);
}

// ---------------------------------------------------------------------------------------------
// Launchers
static inline unsigned grid_for(size_t n_threads, int block) {
    const size_t b = (n_threads + block - 1) / block;
    return (unsigned)((b + kXcds - 1) / kXcds * kXcds);  // multiple of the XCD count (see xcd_block)
}

#define LAUNCH_CHECK()                       \
    do {                                     \
        hipError_t _e = hipGetLastError();   \
        if (_e != hipSuccess) return _e;     \
    } while (0)
#define MARK(i)                                         \
    do {                                                \
        if (ev) {                                       \
            hipError_t _r = hipEventRecord(ev[i], s);   \
            if (_r != hipSuccess) return _r;            \
        }                                               \
    } while (0)

#if !defined(FESOMTRA_INST) || (FESOMTRA_INST & 1)
// Step counter of the stochastic rounding: set to `v`, or advanced by one (skipping 0, which
// means "round to nearest"), by a single thread, so a captured graph advances it at every replay.
__global__ void k_seed(uint32_t *c, uint32_t v, int advance) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t x = advance ? c[0] + 1u : v;
        if (advance && x == 0u) x = 1u;
        c[0] = x;
    }
}
hipError_t launch_seed(uint32_t *c, uint32_t v, int advance, hipStream_t s) {
    k_seed<<<1, 64, 0, s>>>(c, v, advance);
    return hipGetLastError();
}
#endif

template <typename S>
__global__ __launch_bounds__(256) void k_copy(S *__restrict__ dst, const S *__restrict__ src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int HOR, bool FCT, int KB, bool GEN, typename S, typename M, typename C>
static void edge_kernel(dim3 g, hipStream_t s, const MeshCol<M> &m, const StepPtrs<S, C> &p, C ph, int lc) {
    if constexpr (KB <= 4) {
        if (m.et_n > 0 && m.et_kb == KB) {  // edge tiles built for this block width
            const dim3 gt(grid_for((size_t)m.et_n, 1), g.y);
            if (p.TAB == p.T) k_edge<true, HOR, FCT, KB, S, M, C, true, GEN><<<gt, 256, 0, s>>>(m, p, ph, lc);
            else k_edge<true, HOR, FCT, KB, S, M, C, false, GEN><<<gt, 256, 0, s>>>(m, p, ph, lc);
            return;
        }
    }
    k_edge<false, HOR, FCT, KB, S, M, C, false, GEN><<<g, 256, 0, s>>>(m, p, ph, lc);
}

template <int KB, bool GEN, typename S, typename M, typename C>
static hipError_t step_kb(const MeshCol<M> &m, const StepPtrs<S, C> &p, const SchemeArgs &sc,
                          const LaunchCfg &cfg, hipStream_t s, hipEvent_t *ev) {
    const int L = m.L;
    auto grid = [&](int n_cols, int lc) {
        return dim3(grid_for((size_t)n_cols * KB, 256), (unsigned)((L + lc - 1) / lc));
    };
    const dim3 ge = grid(m.E, cfg.lc_edge), gl = grid(m.N, cfg.lc_lo), gb = grid(m.N, cfg.lc_bounds),
               gf = grid(m.N, cfg.lc_final);
    const C ph = (C)sc.ph, pv = (C)sc.pv;
    // halo-tiled node kernels when the tiles were built for this block width
    bool tiled = false;
    if constexpr (KB <= 4) tiled = m.ht_n > 0 && m.ht_kb == KB;
    const unsigned gt = grid_for((size_t)m.ht_n, 1);
    auto tgrid = [&](int lc) { return dim3(gt, (unsigned)((L + lc - 1) / lc)); };
    // The low-order solution (with the implicit split and the cavity rows) reads only the
    // transport and the values, not the antidiffusive fluxes: with the limiter it runs on the aux
    // stream beside the gradient and edge-flux kernels, joined before the bounds kernel (not while
    // the stages are timed one by one).
    auto lo_group = [&](hipStream_t st) -> hipError_t {
        if (tiled) {
            if constexpr (KB <= 4) {
                const dim3 tl = tgrid(cfg.lc_lo);
                switch (sc.ver) {
                case VER_UPW1: k_lo<true, VER_UPW1, KB, S, M, C, GEN><<<tl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
                case VER_QR4C: k_lo<true, VER_QR4C, KB, S, M, C, GEN><<<tl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
                case VER_CDIFF: k_lo<true, VER_CDIFF, KB, S, M, C, GEN><<<tl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
                case VER_PPM: k_lo<true, VER_PPM, KB, S, M, C, GEN><<<tl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
                default: return hipErrorInvalidValue;
                }
            }
        } else {
            switch (sc.ver) {
            case VER_UPW1: k_lo<false, VER_UPW1, KB, S, M, C, GEN><<<gl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
            case VER_QR4C: k_lo<false, VER_QR4C, KB, S, M, C, GEN><<<gl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
            case VER_CDIFF: k_lo<false, VER_CDIFF, KB, S, M, C, GEN><<<gl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
            case VER_PPM: k_lo<false, VER_PPM, KB, S, M, C, GEN><<<gl, 256, 0, st>>>(m, p, pv, sc.dt, cfg.lc_lo, sc.wsplit); break;
            default: return hipErrorInvalidValue;
            }
        }
        LAUNCH_CHECK();
        if (sc.wsplit) {
            if (!p.tri) return hipErrorInvalidValue;
            k_vimpl<KB, S, M, C, GEN><<<dim3(grid_for((size_t)m.N * KB, 256)), 256, 0, st>>>(m, p, sc.dt);
            LAUNCH_CHECK();
        }
        if (m.cq_n > 0) {
            k_cavity_aux<KB><<<dim3(grid_for((size_t)m.cq_n * KB, 256)), 256, 0, st>>>(m, p, sc.cavity_fill ? 1 : 0);
            LAUNCH_CHECK();
        }
        return hipSuccess;
    };
    const bool conc = sc.fct && cfg.aux && cfg.fork && cfg.join && !ev;
    if (conc) {
        hipError_t e = hipEventRecord(cfg.fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(cfg.aux, cfg.fork, 0);
        if (e == hipSuccess) e = lo_group(cfg.aux);
        if (e == hipSuccess) e = hipEventRecord(cfg.join, cfg.aux);
        if (e != hipSuccess) return e;
    }
    MARK(0);
    if (sc.hor != HOR_UPW1 && p.mg && m.mg_n > 0) {
        k_mgrad<KB><<<dim3(grid_for((size_t)m.mg_n * KB, 256)), 256, 0, s>>>(m, p);
        LAUNCH_CHECK();
    }
    switch (sc.hor * 2 + (sc.fct ? 1 : 0)) {
    case HOR_UPW1 * 2 + 0: edge_kernel<HOR_UPW1, false, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    case HOR_UPW1 * 2 + 1: edge_kernel<HOR_UPW1, true, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    case HOR_MUSCL * 2 + 0: edge_kernel<HOR_MUSCL, false, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    case HOR_MUSCL * 2 + 1: edge_kernel<HOR_MUSCL, true, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    case HOR_MFCT * 2 + 0: edge_kernel<HOR_MFCT, false, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    case HOR_MFCT * 2 + 1: edge_kernel<HOR_MFCT, true, KB, GEN>(ge, s, m, p, ph, cfg.lc_edge); break;
    default: return hipErrorInvalidValue;
    }
    LAUNCH_CHECK();
    MARK(1);
    if (sc.fct) {
        if (conc) {
            const hipError_t e = hipStreamWaitEvent(s, cfg.join, 0);
            if (e != hipSuccess) return e;
        } else {
            const hipError_t e = lo_group(s);
            if (e != hipSuccess) return e;
        }
        if (tiled) {
            if constexpr (KB <= 4) {
                // the fused limiter where the caller asked for it: FP64 with the two-ring tiles, an
                // output buffer apart from the values (or the drop-in's accumulators)
                bool fused_lim = false;
                if constexpr (std::is_same<C, double>::value && std::is_same<S, double>::value && !GEN)
                    fused_lim = m.rt_n > 0 && !p.ab_next && (p.Tout || p.acc_h);
                if (p.Tout && !fused_lim) return hipErrorInvalidValue;  // the update must reach Tout
                MARK(2);
                if (fused_lim) {
                    if constexpr (std::is_same<C, double>::value && std::is_same<S, double>::value && !GEN) {
                        MARK(3);
                        k_limit<KB, S, M, C><<<dim3(grid_for((size_t)m.rt_n, 1)), 256, 0, s>>>(m, p, sc.dt);
                    }
                } else {
                    k_bounds<true, KB, S, M, C, GEN><<<tgrid(cfg.lc_bounds), 256, 0, s>>>(m, p, sc.dt, cfg.lc_bounds);
                    LAUNCH_CHECK();
                    MARK(3);
                    if (p.ab_next) k_final<true, true, KB, S, M, C, true, GEN><<<tgrid(cfg.lc_final), 256, 0, s>>>(m, p, sc.dt, cfg.lc_final);
                    else k_final<true, true, KB, S, M, C, false, GEN><<<tgrid(cfg.lc_final), 256, 0, s>>>(m, p, sc.dt, cfg.lc_final);
                }
            }
        } else {
            MARK(2);
            k_bounds<false, KB, S, M, C, GEN><<<gb, 256, 0, s>>>(m, p, sc.dt, cfg.lc_bounds);
            LAUNCH_CHECK();
            MARK(3);
            if (p.ab_next) k_final<false, true, KB, S, M, C, true, GEN><<<gf, 256, 0, s>>>(m, p, sc.dt, cfg.lc_final);
            else if (KB == 64 && !p.acc_h && !p.dbg_h && !p.dbg_v)  // no accumulators or kept fluxes
                k_final<false, true, KB, S, M, C, false, GEN, true><<<gf, 256, 0, s>>>(m, p, sc.dt, cfg.lc_final);
            else k_final<false, true, KB, S, M, C, false, GEN><<<gf, 256, 0, s>>>(m, p, sc.dt, cfg.lc_final);
        }
    } else {
        switch (sc.ver) {
        case VER_UPW1: k_vert<VER_UPW1, KB, S, M, C, GEN><<<gl, 256, 0, s>>>(m, p, pv, sc.dt, cfg.lc_lo); break;
        case VER_QR4C: k_vert<VER_QR4C, KB, S, M, C, GEN><<<gl, 256, 0, s>>>(m, p, pv, sc.dt, cfg.lc_lo); break;
        case VER_CDIFF: k_vert<VER_CDIFF, KB, S, M, C, GEN><<<gl, 256, 0, s>>>(m, p, pv, sc.dt, cfg.lc_lo); break;
        case VER_PPM: k_vert<VER_PPM, KB, S, M, C, GEN><<<gl, 256, 0, s>>>(m, p, pv, sc.dt, cfg.lc_lo); break;
        default: return hipErrorInvalidValue;
        }
        LAUNCH_CHECK();
        MARK(2);
        MARK(3);
        // single precision without limiter: the update adds the upwind part of valuesAB at the
        // far nodes, so when valuesAB is values and the update is in place it must read a snapshot
        // taken before the first store (the LO buffer, unused without the limiter, holds it)
        StepPtrs<S, C> pf = p;
        if constexpr (!std::is_same<C, double>::value) {
            if (p.TAB == p.T && !p.acc_h) {
                const size_t n = (size_t)L * m.Np * KB;
                static_assert(sizeof(S) <= sizeof(C), "the snapshot fits the LO buffer");
                k_copy<S><<<dim3(grid_for(n, 256)), 256, 0, s>>>(reinterpret_cast<S *>(p.LO), p.T, n);
                LAUNCH_CHECK();
                pf.TAB = reinterpret_cast<const S *>(p.LO);
            }
        }
        if (pf.ab_next) k_final<false, false, KB, S, M, C, true, GEN><<<gf, 256, 0, s>>>(m, pf, sc.dt, cfg.lc_final);
        else k_final<false, false, KB, S, M, C, false, GEN><<<gf, 256, 0, s>>>(m, pf, sc.dt, cfg.lc_final);
    }
    LAUNCH_CHECK();
    MARK(4);
    return hipSuccess;
}

template <typename M>
hipError_t launch_qflux(const MeshCol<M> &m, hipStream_t s) {
    if (m.L < 1 || m.E < 1) return hipSuccess;
    if (m.nmem > 1) k_qflux<M, true><<<dim3(grid_for((size_t)m.E * m.nmem, 256)), 256, 0, s>>>(m);
    else k_qflux<M><<<dim3(grid_for((size_t)m.E, 256)), 256, 0, s>>>(m);
    return hipGetLastError();
}

template <typename S, typename M, typename C>
hipError_t launch_step(const MeshCol<M> &m, const StepPtrs<S, C> &p, int kb, const SchemeArgs &sc,
                       const LaunchCfg &cfg, hipStream_t s, hipEvent_t *ev) {
    if (cfg.lc_edge < 1 || cfg.lc_lo < 1 || cfg.lc_bounds < 1 || cfg.lc_final < 1) return hipErrorInvalidValue;
    if (kb < 64 && (p.mask >> kb) != 0ull) return hipErrorInvalidValue;  // lanes beyond the block
    if (m.cq_n > 0 && sc.fct && !p.cqv) return hipErrorInvalidValue;    // cavity bounds need their rows
    // per-node time steps (m.dtn), cavities (m.ulev_n) or ensemble members select the general
    // kernels; the others keep dt a scalar, every top at level 1 and one velocity
    const bool ndt = m.dtn != nullptr || m.ulev_n != nullptr || m.nmem > 1;
    if (m.nmem > 1 && !p.lane_mem) return hipErrorInvalidValue;  // members need the lane table
    if (p.Tout && (!sc.fct || kb > 4 || m.ht_kb != kb || m.rt_n < 1)) return hipErrorInvalidValue;
    switch (kb) {
    case 1: return ndt ? step_kb<1, true>(m, p, sc, cfg, s, ev) : step_kb<1, false>(m, p, sc, cfg, s, ev);
    case 2: return ndt ? step_kb<2, true>(m, p, sc, cfg, s, ev) : step_kb<2, false>(m, p, sc, cfg, s, ev);
    case 4: return ndt ? step_kb<4, true>(m, p, sc, cfg, s, ev) : step_kb<4, false>(m, p, sc, cfg, s, ev);
    case 8: return ndt ? step_kb<8, true>(m, p, sc, cfg, s, ev) : step_kb<8, false>(m, p, sc, cfg, s, ev);
    case 16: return ndt ? step_kb<16, true>(m, p, sc, cfg, s, ev) : step_kb<16, false>(m, p, sc, cfg, s, ev);
    case 32: return ndt ? step_kb<32, true>(m, p, sc, cfg, s, ev) : step_kb<32, false>(m, p, sc, cfg, s, ev);
    case 64: return ndt ? step_kb<64, true>(m, p, sc, cfg, s, ev) : step_kb<64, false>(m, p, sc, cfg, s, ev);
    default: return hipErrorInvalidValue;
    }
}

template <typename S>
hipError_t launch_relayout(const double *ref_in, S *dev_out, const S *dev_in, double *ref_out,
                           const int *perm, int N, int L, size_t Np, int kb, int lane, int dir, hipStream_t s,
                           float *res) {
    const size_t total = (size_t)N * L;
    if (total == 0) return hipSuccess;
    if (lane < 0 || lane >= kb) return hipErrorInvalidValue;
    const dim3 g((unsigned)((N + kRelayNodes - 1) / kRelayNodes), (unsigned)((L + kRelayLevels - 1) / kRelayLevels));
    k_relayout<S><<<g, 256, 0, s>>>(ref_in, dev_out, dev_in, ref_out, perm, N, L, Np, kb, lane, dir, res);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_ab2(const S *T, S *AB, S *old, size_t n, double eps, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)kXcds * 2048);
    k_ab2<S><<<dim3((unsigned)blocks), 256, 0, s>>>(T, AB, old, n, 0.5 + eps, 1.5 + eps);
    return hipGetLastError();
}

template <typename S, typename M>
hipError_t launch_stats(const MeshCol<M> &m, const S *T, int kb, int lane, int all_levels, double4 *part,
                        int nblocks, hipStream_t s, const float *res) {
    if (lane < 0 || lane >= kb) return hipErrorInvalidValue;
    k_stats_partial<S, M><<<nblocks, 256, 0, s>>>(m, T, res, kb, lane, all_levels, part);
    return hipGetLastError();
}

// Instantiations: FP64 (bit-exact path) and FP32 reconstruction arithmetic with FP32 / FP16 /
// BF16 storage (FP64 mesh and velocities in every mode).
// FESOMTRA_INST selects a subset (bit 0: FP64, 1: FP32, 2: FP16 storage, 3: BF16 storage) so the
// four families can be compiled as separate objects in parallel.
#ifndef FESOMTRA_INST
#define FESOMTRA_INST 15
#endif
#define FESOMTRA_INSTANTIATE(S_, M_, C_)                                                                    \
    template hipError_t launch_step<S_, M_, C_>(const MeshCol<M_> &, const StepPtrs<S_, C_> &, int,         \
                                                const SchemeArgs &, const LaunchCfg &, hipStream_t,         \
                                                hipEvent_t *);                                              \
    template hipError_t launch_relayout<S_>(const double *, S_ *, const S_ *, double *, const int *, int, \
                                            int, size_t, int, int, int, hipStream_t, float *);             \
    template hipError_t launch_stats<S_, M_>(const MeshCol<M_> &, const S_ *, int, int, int, double4 *,    \
                                             int, hipStream_t, const float *);                               \
    template hipError_t launch_ab2<S_>(const S_ *, S_ *, S_ *, size_t, double, hipStream_t);
#if FESOMTRA_INST & 1
FESOMTRA_INSTANTIATE(double, double, double)
template hipError_t launch_qflux<double>(const MeshCol<double> &, hipStream_t);
#endif
#if FESOMTRA_INST & 2
FESOMTRA_INSTANTIATE(float, double, float)
#endif
#if FESOMTRA_INST & 4
FESOMTRA_INSTANTIATE(f16_t, double, float)
#endif
#if FESOMTRA_INST & 8
FESOMTRA_INSTANTIATE(bf16_t, double, float)
#endif

}  // namespace fesomtra
