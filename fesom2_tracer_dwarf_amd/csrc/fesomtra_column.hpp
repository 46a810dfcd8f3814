// Per-column formulas of the tracer advection step (device functions).
//
// Every expression keeps the operation order of the reference routine it cites so that the
// double instantiation reproduces the serial reference bit for bit.  Levels are 1-based (nz),
// interfaces ulevels_nod2D..nlevels_nod2D (the top is 1 without ice-shelf cavities).
#pragma once

#include "fesomtra_types.hpp"

namespace fesomtra {

__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
// Column owned by this thread; consecutive columns of one XCD stay contiguous in the
// space-filling-curve order (workgroups are dispatched round-robin over the XCDs).
__device__ __forceinline__ int column_index() { return xcd_block() * blockDim.x + threadIdx.x; }
// (column, lane) of this thread for a tracer block of width KB (a power of two).  With KB = 64 a
// wavefront is exactly one column: the column index is made wavefront-uniform, so every
// per-column quantity lives in scalar registers and its loads are scalar loads.
template <int KB>
__device__ __forceinline__ void column_lane(int &col, int &lane) {
    const int g = column_index();
    if constexpr (KB == 64) {
        col = __builtin_amdgcn_readfirstlane(g / KB);
        lane = (int)(threadIdx.x & 63u);
    } else {
        col = g / KB;
        lane = g % KB;
    }
}
__device__ __forceinline__ bool lane_on(unsigned long long mask, int lane) { return (mask >> lane) & 1ull; }

// Reads lane `lane` of a tracer block: ld(cell) = block[cell * KB + lane] in compute type C.
template <int KB, typename C, typename S>
struct LaneLd {
    const S *a;
    int lane;
    __device__ __forceinline__ C operator()(size_t cell) const { return ldc<C>(a, cell * KB + lane); }
};

// First-order upwind edge flux (lib/oce_adv_tra_hor.F90:178-181).
template <typename C>
__device__ __forceinline__ C upw1_edge(C t1, C t2, C v) {
    return -C(0.5) * (t1 * (v + fabs(v)) + t2 * (v - fabs(v)));
}

// tr_xy(comp, nz, el) = sum over the element's nodes of gradient_sca * T, in node order.  In
// single precision (C = float) it is evaluated as g1 (T_b - T_a) + g2 (T_c - T_a), equal since
// the three coefficients of a P1 gradient sum to zero, and exactly zero for a uniform field.
template <typename C, typename LD>
__device__ __forceinline__ C tr_xy(const C g[3], const LD &t, size_t off, int a, int b, int c) {
    if constexpr (std::is_same<C, double>::value) {
        return g[0] * t(off + a) + g[1] * t(off + b) + g[2] * t(off + c);
    } else {
        const C ta = t(off + a);
        return g[1] * (t(off + b) - ta) + g[2] * (t(off + c) - ta);
    }
}

// Area-weighted mean element gradient around node n at level nz over the elements wet there
// (lib/oce_muscl_adv.F90:394-409), accumulated in nod_in_elem2D order.
template <typename C, typename LD, typename M>
__device__ __forceinline__ void node_mean_grad(const MeshCol<M> &m, const LD &t, int n, int nz, C &gx,
                                               C &gy) {
    C tvol = 0, tx = 0, ty = 0;
    const size_t off = (size_t)(nz - 1) * m.Np;
    const int cnt = cld(m.nie_num, n);
    for (int i = 0; i < cnt; ++i) {
        const int el = cld(m.nie, (size_t)n * m.max_adj + i);
        if (cld(m.nlev_e, el) - 1 < nz || nz < elem_top(m, el)) continue;
        const C ar = (C)cld(m.elem_area, el);
        const int en[3] = {cld(m.enodes, (size_t)el * 3), cld(m.enodes, (size_t)el * 3 + 1),
                           cld(m.enodes, (size_t)el * 3 + 2)};
        C g0[3], g1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g0[c] = (C)cld(m.gsca, (size_t)el * 6 + c);
            g1[c] = (C)cld(m.gsca, (size_t)el * 6 + 3 + c);
        }
        tvol = tvol + ar;
        tx = tx + tr_xy<C>(g0, t, off, en[0], en[1], en[2]) * ar;
        ty = ty + tr_xy<C>(g1, t, off, en[0], en[1], en[2]) * ar;
    }
    gx = tx / tvol;
    gy = ty / tvol;
}

// MUSCL / MFCT high-order flux minus fin (lib/oce_adv_tra_hor.F90:368-382, 748-790);
// c1, c2 are MUSCL's c_lo switches.
template <int HOR, typename C>
__device__ __forceinline__ C ho_edge(C t1, C t2, const C gr[4], C v, C dx, C dy, C a, C num_ord,
                                     C c1, C c2, C fin) {
    const C r = (C)kREarth;
    C Tmean1, Tmean2;
    if (HOR == HOR_MUSCL) {
        Tmean2 = t2 - (C(2) * (t2 - t1) + dx * a * gr[1] + dy * r * gr[3]) / C(6) * c2;
        Tmean1 = t1 + (C(2) * (t2 - t1) + dx * a * gr[0] + dy * r * gr[2]) / C(6) * c1;
    } else {
        Tmean2 = t2 - (C(2) * (t2 - t1) + dx * a * gr[1] + dy * r * gr[3]) / C(6);
        Tmean1 = t1 + (C(2) * (t2 - t1) + dx * a * gr[0] + dy * r * gr[2]) / C(6);
    }
    const C cHO = (v + fabs(v)) * Tmean1 + (v - fabs(v)) * Tmean2;
    return -C(0.5) * (C(1) - num_ord) * cHO - v * num_ord * C(0.5) * (Tmean1 + Tmean2) - fin;
}

// MUSCL / MFCT high-order flux minus the first-order upwind flux of the same values, in the
// difference form  -(1-th)/2 [(v+|v|) d1 + (v-|v|) d2] - th [v (d1+d2)/2 - |v| (a1-a2)/2]  with
// d1 = Tmean1 - a1, d2 = Tmean2 - a2: algebraically the reference's ho_edge(...) - upw1_edge(a1,
// a2, v), evaluated so that only small differences are rounded (used when C is single
// precision).
template <int HOR, typename C>
__device__ __forceinline__ C ho_minus_lo(C a1, C a2, const C gr[4], C v, C dx, C dy, C a, C th,
                                         C c1, C c2) {
    const C r = (C)kREarth, sixth = C(1) / C(6);  // single precision: multiply by 1/6
    C d1 = (C(2) * (a2 - a1) + dx * a * gr[0] + dy * r * gr[2]) * sixth;
    C d2 = -(C(2) * (a2 - a1) + dx * a * gr[1] + dy * r * gr[3]) * sixth;
    if (HOR == HOR_MUSCL) {
        d1 = d1 * c1;
        d2 = d2 * c2;
    }
    const C av = fabs(v);
    return -C(0.5) * (C(1) - th) * ((v + av) * d1 + (v - av) * d2) -
           th * (v * C(0.5) * (d1 + d2) - C(0.5) * av * (a1 - a2));
}

// ------------------------------------------------------------------------------------------
// Vertical fluxes (lib/oce_adv_tra_ver.F90).  A Col reads one node column straight from memory
// (used by the less common CDIFF / PPM schemes); QR4C and UPW1 use register windows instead.
template <typename C, typename S, typename M, int KB>
struct Col {
    const S *t;
    const M *w, *ar, *hn, *ho;
    size_t Np;
    int n, lane, nzmin, nzmax;
    Member mb;  // the lane's velocity member (w is member-interleaved)
    __device__ __forceinline__ C T(int nz) const { return ldc<C>(t, ((size_t)(nz - 1) * Np + n) * KB + lane); }
    __device__ __forceinline__ C W(int j) const { return (C)cld(w, mb.at((size_t)(j - 1) * Np + n)); }
    __device__ __forceinline__ C A(int j) const { return (C)cld(ar, (size_t)(j - 1) * Np + n); }
    __device__ __forceinline__ C HN(int nz) const { return (C)cld(hn, (size_t)(nz - 1) * Np + n); }
    __device__ __forceinline__ C HO(int nz) const { return (C)cld(ho, (size_t)(nz - 1) * Np + n); }
};

// adv_tra_ver_cdiff (:687-708) at interface j, minus fin; the bottom interface keeps fin.
template <typename C, typename S, typename M, int KB>
__device__ __forceinline__ C ver_cdiff(const Col<C, S, M, KB> &c, int j, C fin) {
    if (j == c.nzmax) return fin;
    if (j == c.nzmin) return -c.W(j) * c.T(j) * c.A(j) - fin;
    const C tv = C(0.5) * (c.T(j - 1) + c.T(j));
    return -tv * c.W(j) * c.A(j) - fin;
}

// PPM interface value tv(j) after all of the reference's assignments (:510-594).
template <typename C, typename S, typename M, int KB>
__device__ __forceinline__ C ppm_tv(const Col<C, S, M, KB> &c, int j) {
    const int nzmin = c.nzmin, nzmax = c.nzmax;
    if (j >= nzmin + 2 && j <= nzmax - 2) {
        const int nz = j - 1;
        const C dzjm1 = c.HN(nz - 1), dzj = c.HN(nz), dzjp1 = c.HN(nz + 1), dzjp2 = c.HN(nz + 2);
        const C tm1 = c.T(nz - 1), t0 = c.T(nz), tp1 = c.T(nz + 1), tp2 = c.T(nz + 2);
        C deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                   ((C(2) * dzjm1 + dzj) / (dzjp1 + dzj) * (tp1 - t0) +
                    (dzj + C(2) * dzjp1) / (dzjm1 + dzj) * (t0 - tm1));
        C deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                     ((C(2) * dzj + dzjp1) / (dzjp2 + dzjp1) * (tp2 - tp1) +
                      (dzjp1 + C(2) * dzjp2) / (dzj + dzjp1) * (tp1 - t0));
        if ((tp1 - t0) * (t0 - tm1) > C(0))
            deltaj = cmin(cmin(C(fabs(deltaj)), C(2) * fabs(tp1 - t0)), C(2) * fabs(t0 - tm1)) * csign1(deltaj);
        else
            deltaj = C(0);
        if ((tp2 - tp1) * (tp1 - t0) > C(0))
            deltajp1 = cmin(cmin(C(fabs(deltajp1)), C(2) * fabs(tp2 - tp1)), C(2) * fabs(tp1 - t0)) * csign1(deltajp1);
        else
            deltajp1 = C(0);
        return t0 + dzj / (dzj + dzjp1) * (tp1 - t0) +
               C(1) / (dzjm1 + dzj + dzjp1 + dzjp2) *
                   ((C(2) * dzjp1 * dzj) / (dzj + dzjp1) *
                        ((dzjm1 + dzj) / (C(2) * dzj + dzjp1) - (dzjp2 + dzjp1) / (C(2) * dzjp1 + dzj)) *
                        (tp1 - t0) -
                    dzj * (dzjm1 + dzj) / (C(2) * dzj + dzjp1) * deltajp1 +
                    dzjp1 * (dzjp1 + dzjp2) / (dzj + C(2) * dzjp1) * deltaj);
    }
    if (j == nzmax) return c.T(nzmax - 1);
    if (j == nzmax - 1) {
        const C s = csign1(c.W(nzmax - 1));
        return -c.T(nzmax - 2) * cmin(s, C(0)) + c.T(nzmax - 1) * cmax(s, C(0));
    }
    if (j == nzmin + 1) return C(0.5) * (c.T(nzmin) + c.T(nzmin + 1));
    return c.T(nzmin);
}

// Limited parabola of layer nz (:601-617).
template <typename C, typename S, typename M, int KB>
__device__ __forceinline__ void ppm_layer(const Col<C, S, M, KB> &c, int nz, C &aL, C &aR, C &aj) {
    const C t = c.T(nz);
    aL = ppm_tv(c, nz);
    aR = ppm_tv(c, nz + 1);
    if ((aR - t) * (t - aL) <= C(0)) {
        aL = t;
        aR = t;
    }
    if ((aR - aL) * (t - C(0.5) * (aL + aR)) > (aR - aL) * (aR - aL) / C(6)) aL = C(3) * t - C(2) * aR;
    if ((aR - aL) * (t - C(0.5) * (aR + aL)) < -(aR - aL) * (aR - aL) / C(6)) aR = C(3) * t - C(2) * aL;
    aj = C(6) * (t - C(0.5) * (aL + aR));
}

// adv_tra_vert_ppm (:596-639) at interface j, minus fin.
template <typename C, typename S, typename M, int KB>
__device__ __forceinline__ C ver_ppm(const Col<C, S, M, KB> &c, int j, C dt, C fin) {
    if (j == c.nzmin) return -ppm_tv(c, c.nzmin) * c.W(c.nzmin) * c.A(c.nzmin) - fin;
    if (j == c.nzmax) return C(0) - fin;
    const C wj = c.W(j);
    C tvert = C(0), aL, aR, aj;
    if (wj > C(0)) {  // "W(nz) > 0" branch of layer j
        ppm_layer(c, j, aL, aR, aj);
        const C x = cmin(wj * dt / c.HO(j), C(1));
        tvert = (-aL - C(0.5) * x * (aR - aL + (C(1) - C(2) / C(3) * x) * aj));
        tvert = tvert * c.A(j) * wj;
    } else if (wj < C(0)) {  // "W(nz+1) < 0" branch of layer j-1
        ppm_layer(c, j - 1, aL, aR, aj);
        const C x = cmin(-wj * dt / c.HO(j - 1), C(1));
        tvert = (-aR + C(0.5) * x * (aR - aL - (C(1) - C(2) / C(3) * x) * aj));
        tvert = tvert * c.A(j) * wj;
    }
    return tvert - fin;
}

// QR4C (:394-436) at interface j of a column with top layer `top` from the register window
// tw[i] = T(j-2+i), i = 0..3 and z[i] = Z_3d_n(j-2+i), minus fin.  A 2-layer column updates its
// middle interface twice, as the reference's statement sequence does.
template <typename C>
__device__ __forceinline__ C ver_qr4c_w(int j, int top, int nzmax, const C tw[4], const C z[4], C zb, C w,
                                        C ar, C num_ord, C fin, const C *qw = nullptr) {
    if (j == top) return -tw[2] * w * ar - fin;
    if (j == nzmax) return C(0) - fin;
    const bool second = (j == top + 1), penult = (j == nzmax - 1);
    if (second || penult) {
        const C cv = -C(0.5) * (tw[1] + tw[2]) * w * ar;
        C f = fin;
        if (second) f = cv - f;
        if (penult) f = cv - f;
        return f;
    }
    // qd, qc, qu: the quotient (T(i) - T(i+1)) / (Z(i) - Z(i+1)) for i = j-2, j-1, j; a caller
    // sweeping the column passes them in (qw), each computed once (same operands, same value)
    const C qc = qw ? qw[1] : (tw[1] - tw[2]) / (z[1] - z[2]);
    const C qu = qw ? qw[2] : (tw[2] - tw[3]) / (z[2] - z[3]);
    const C qd = qw ? qw[0] : (tw[0] - tw[1]) / (z[0] - z[1]);
    const C Tmean1 = tw[2] + (C(2) * qc + qu) * (zb - z[2]) / C(3);
    const C Tmean2 = tw[1] + (C(2) * qc + qd) * (zb - z[1]) / C(3);
    const C Tmean = (w + fabs(w)) * Tmean1 + (w - fabs(w)) * Tmean2;
    return (-C(0.5) * (C(1) - num_ord) * Tmean - num_ord * (C(0.5) * (Tmean1 + Tmean2)) * w) * ar - fin;
}

// adv_tra_ver_upw1 (:300-327) at interface j of a column with top layer `top`, with T(j-1) = tm,
// T(j) = tj, minus fin.
template <typename C>
__device__ __forceinline__ C ver_upw1_w(int j, int top, int nzmax, C tm, C tj, C w, C ar, C fin) {
    if (j == top) return -w * tj * ar - fin;
    if (j == nzmax) return C(0) - fin;
    return -C(0.5) * (tj * (w + fabs(w)) + tm * (w - fabs(w))) * ar - fin;
}

// QR4C's three quotients at interface j, (T(i) - T(i+1)) / (Z(i) - Z(i+1)) for i = j-2, j-1, j,
// kept while a column is swept downward: one new quotient per interface.
template <typename C>
struct QuotWin {
    C q[3];
    __device__ __forceinline__ static C quot(C ta, C tb, C za, C zb) { return (ta - tb) / (za - zb); }
    // all three for the window tw = TAB(j-2..j+1), z = Z(j-2..j+1)
    __device__ __forceinline__ void init(const C tw[4], const C z[4]) {
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = quot(tw[i], tw[i + 1], z[i], z[i + 1]);
    }
    // after the window moved down by one interface
    __device__ __forceinline__ void slide(const C tw[4], const C z[4]) {
        q[0] = q[1];
        q[1] = q[2];
        q[2] = quot(tw[2], tw[3], z[2], z[3]);
    }
};

// Column readers of one node (1-based levels; 0 outside the wet range).
template <typename C, typename S, typename M, int KB>
struct NodeCol {
    size_t Np;
    int n, lane, L;
    Member mb;  // the lane's velocity member
    // velocity-dependent node-interface array (w, w_e, w_i) of the lane's member at interface j
    __device__ __forceinline__ C vel(const M *a, int j) const { return (C)cld(a, mb.at((size_t)(j - 1) * Np + n)); }
    __device__ __forceinline__ C lay(const S *a, int nz) const {
        return (nz >= 1 && nz <= L) ? ldc<C>(a, ((size_t)(nz - 1) * Np + n) * KB + lane) : C(0);
    }
    __device__ __forceinline__ C mlay(const M *a, int nz) const {
        return (nz >= 1 && nz <= L) ? (C)cld(a, (size_t)(nz - 1) * Np + n) : C(0);
    }
    __device__ __forceinline__ C itf(const M *a, int j) const { return (C)cld(a, (size_t)(j - 1) * Np + n); }
};

// High-order vertical flux at interface j minus fin for the window tw = TAB(j-2..j+1),
// z = Z(j-2..j+1) (QR4C: with the window's quotients qw when given, see QuotWin).
template <int VER, typename C, typename S, typename M, int KB>
__device__ __forceinline__ C ver_ho_at(int j, int nzmax, const C tw[4], const C z[4], C zbj, C wj, C aj,
                                       C num_ord, C dt, const Col<C, S, M, KB> &col, C fin,
                                       const C *qw = nullptr) {
    if (VER == VER_QR4C) return ver_qr4c_w<C>(j, col.nzmin, nzmax, tw, z, zbj, wj, aj, num_ord, fin, qw);
    if (VER == VER_UPW1) return ver_upw1_w<C>(j, col.nzmin, nzmax, tw[1], tw[2], wj, aj, fin);
    if (VER == VER_CDIFF) return ver_cdiff(col, j, fin);
    return ver_ppm(col, j, dt, fin);
}

template <typename C, typename S, typename M, int KB>
__device__ __forceinline__ Col<C, S, M, KB> make_col(const S *t, const M *w, const MeshCol<M> &m, int n,
                                                     int lane, int nzmax, int top = 1, Member mb = Member{0, 1}) {
    Col<C, S, M, KB> c;
    c.mb = mb;
    c.t = t;
    c.w = w;
    c.ar = m.area;
    c.hn = m.hnode_new;
    c.ho = m.hnode;
    c.Np = m.Np;
    c.n = n;
    c.lane = lane;
    c.nzmin = top;
    c.nzmax = nzmax;
    return c;
}

}  // namespace fesomtra
