// Fused single-pass FCT step over tiles of nodes (one workgroup per tile, one thread per node of
// the tile's region), FP64, one tracer per pass.
//
// A tile owns a stretch of the Hilbert-ordered nodes (ring 0); its region adds the three rings
// of neighbours around it.  Walking the levels top-down, each level a runs four stages separated
// by workgroup barriers, exchanging values between the region's nodes through LDS:
//   1  T(a+1) of every region node -> LDS
//   2  element gradients at level a -> LDS (MUSCL / MFCT); low-order solution LO(a+1) of rings
//      0-2 (edge upwind fluxes summed in reference edge order, vertical upwind flux) -> LDS, with
//      the vertical antidiffusive flux of the column kept in registers
//   3  rings 0-1: cluster bounds of layer a+1, antidiffusive edge fluxes of layer a (kept in
//      registers), P+/P-, Q+/Q-, limiter ratios R+/R-(a) -> LDS
//   4  ring 0: limiting, flux divergence and update of layer a-1 into the output buffer
// so the only HBM traffic per node-layer is the tracer in and out and the mesh / velocity data;
// no edge or node intermediates are stored.  Ring 3 only supplies T; ring 2 also LO (needed by the
// cluster bounds of ring 1); ring 1 also R (needed by the limiter of ring 0's edges).
//
// Every value is computed by the same device functions, in the same operation order and with the
// same summation order as the multi-kernel path (lib/oce_adv_tra_driver.F90:102-216,
// lib/oce_adv_tra_hor.F90, lib/oce_adv_tra_ver.F90, lib/oce_adv_tra_fct.F90:96-471,
// lib/oce_adv_tra_driver.F90:298-450, src/fesom_analytic.F90:419-437), so results stay
// bit-identical to the serial FP64 reference.  T is read from one buffer and written to another
// (neighbouring tiles read the old values as their halo).
#include "fesomtra_column.hpp"
#include "fesomtra_tile.hpp"

namespace fesomtra {

template <int HOR, int VER>
__global__ __launch_bounds__(kTileThreads) void k_tile(const MeshCol<double> m, const TileMesh tm,
                                                       const double *__restrict__ Tin,
                                                       double *__restrict__ Tout, double num_ord,
                                                       double pv, double dt) {
    using D = double;
    __shared__ D sT[2][kTileThreads];
    __shared__ D sLO[kTileThreads];
    __shared__ D sRp[2][kTileThreads], sRm[2][kTileThreads];
    __shared__ D sGx[kTileElems], sGy[kTileElems];
    // the tile's tables, staged once
    __shared__ int2 sEnt[kTileEntries];
    __shared__ int sEp[kTileThreads + 1], sElp[kTileThreads + 1];
    __shared__ int sEl[kTileSel];
    __shared__ int sElId[kTileElems], sElLoc[kTileElems];

    const int tile = xcd_block();
    const int slot0 = cld(tm.t_slot0, tile), nslot = cld(tm.t_nslot, tile);
    const int el0 = cld(tm.t_el0, tile), nel = cld(tm.t_nel, tile);
    const int lmax = cld(tm.t_lmax, tile);
    const int r = (int)threadIdx.x;
    {
        const int sp0 = cld(tm.t_sp0, tile), ent0 = cld(tm.t_ent0, tile), nent = cld(tm.t_nent, tile);
        const int sel0 = cld(tm.t_sel0, tile), nsel = cld(tm.t_nsel, tile);
        for (int i = r; i < nent; i += kTileThreads) sEnt[i] = cld(tm.ent, ent0 + i);
        for (int i = r; i <= nslot; i += kTileThreads) {
            sEp[i] = cld(tm.s_eptr, sp0 + i);
            sElp[i] = cld(tm.s_elptr, sp0 + i);
        }
        for (int i = r; i < nsel; i += kTileThreads) sEl[i] = cld(tm.sel, sel0 + i);
        for (int i = r; i < nel; i += kTileThreads) {
            sElId[i] = cld(tm.el_id, el0 + i);
            sElLoc[i] = cld(tm.el_loc, el0 + i);
        }
    }
    __syncthreads();
    const bool act = r < nslot;
    const int s = slot0 + (act ? r : 0);
    const int n = act ? cld(tm.s_node, s) : 0;
    const int ring = act ? cld(tm.s_ring, s) : 4;
    const int nzmax = act ? cld(m.nlev_n, n) : 0, L = nzmax - 1;
    const int dry_from = act ? cld(m.nlev_min, n) : 0;
    const int ep0 = act ? sEp[r] : 0, ep1 = act ? sEp[r + 1] : 0;
    const size_t Np = m.Np, Ep = m.Ep;
    const NodeCol<D, double, double, 1> nc{Np, n, 0, L};
    const Col<D, double, double, 1> col = make_col<D, double, double, 1>(Tin, m.w, m, n, 0, nzmax);

    // own column state (index i of a window holds level base + i)
    D tw[5];      // T(a-1 .. a+3)
    D z[4];       // Z(a .. a+3), QR4C
    D lo[3];      // LO(a-1), LO(a), LO(a+1)
    D cmx[3], cmn[3];  // cluster bounds of layers a-1, a, a+1
    D av[4];      // antidiffusive vertical flux at interfaces a-1 .. a+2
    D rp[3], rm[3];    // R+/R- of layers a-2, a-1, a
    D lo_top = 0;      // low-order vertical flux at interface a+1
    D ah[kTileDeg], ahp[kTileDeg];  // antidiffusive edge fluxes of layers a and a-1
#pragma unroll
    for (int i = 0; i < 5; ++i) tw[i] = nc.lay(Tin, i - 1);  // levels -1..3 (0 outside)
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] = (VER == VER_QR4C) ? nc.mlay(m.Z, i) : D(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) lo[i] = cmx[i] = cmn[i] = rp[i] = rm[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = 0;
#pragma unroll
    for (int i = 0; i < kTileDeg; ++i) ah[i] = ahp[i] = 0;
    // T(1) of the region into the buffer of level 1
    sT[1][r] = act ? tw[2] : D(0);

    for (int a = 0; a <= lmax + 1; ++a) {
        const int cur = a & 1, nxt = (a + 1) & 1;
        // ---- stage 1: T(a+1) -> LDS (T(a+1) = tw[2] after the shift below)
        if (a > 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) tw[i] = tw[i + 1];
            tw[4] = act ? nc.lay(Tin, a + 3) : D(0);
            if (VER == VER_QR4C) {
#pragma unroll
                for (int i = 0; i < 3; ++i) z[i] = z[i + 1];
                z[3] = act ? nc.mlay(m.Z, a + 3) : D(0);
            }
            sT[nxt][r] = act ? tw[2] : D(0);
        }
        // now tw = T(a-1), T(a), T(a+1), T(a+2), T(a+3); z = Z(a .. a+3)
        __syncthreads();
        // ---- stage 2: gradients of level a; LO(a+1) of rings 0-2
        if (HOR != HOR_UPW1 && a >= 1) {
            for (int le = r; le < nel; le += kTileThreads) {
                const int el = sElId[le], pk = sElLoc[le];
                const int l0 = pk & 1023, l1 = (pk >> 10) & 1023, l2 = pk >> 20;
                D g0[3], g1[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    g0[c] = cld(m.gsca, (size_t)el * 6 + c);
                    g1[c] = cld(m.gsca, (size_t)el * 6 + 3 + c);
                }
                auto ld = [&](size_t i) { return sT[cur][i]; };
                sGx[le] = tr_xy<D>(g0, ld, 0, l0, l1, l2);
                sGy[le] = tr_xy<D>(g1, ld, 0, l0, l1, l2);
            }
        }
        {
            const int j = a + 2;  // bottom interface of layer a+1
            D lo_new = 0;
            if (act && ring <= 2 && a + 1 <= L) {
                const bool bottom = (j == nzmax);
                const D wej = nc.itf(m.we, j), wj = nc.itf(m.w, j), aj = nc.itf(m.area, j);
                const D zbj = (VER == VER_QR4C && !bottom) ? nc.itf(m.zbar, j) : D(0);
                if (a == 0) {  // surface interface 1: window T(-1 .. 2) = tw[0..3], Z(-1 .. 2)
                    const D we1 = nc.itf(m.we, 1), w1 = nc.itf(m.w, 1), a1 = nc.itf(m.area, 1);
                    const D zb1 = (VER == VER_QR4C) ? nc.itf(m.zbar, 1) : D(0);
                    const D t1w[4] = {tw[0], tw[1], tw[2], tw[3]};
                    const D z1w[4] = {D(0), z[0], z[1], z[2]};
                    lo_top = ver_upw1_w<D>(1, 1, nzmax, tw[1], tw[2], we1, a1, D(0));
                    av[2] = ver_ho_at<VER>(1, nzmax, t1w, z1w, zb1, w1, a1, pv, dt, col, lo_top);
                }
                // window of interface j: T(j-2 .. j+1) = T(a .. a+3) = tw[1..4]
                const D tj[4] = {tw[1], tw[2], tw[3], tw[4]};
                const D lo_bot = ver_upw1_w<D>(j, 1, nzmax, tw[2], tw[3], wej, aj, D(0));
                av[3] = ver_ho_at<VER>(j, nzmax, tj, z, zbj, wj, aj, pv, dt, col, lo_bot);
                const size_t cell = (size_t)a * Np + n;  // layer a+1
                const double h = cld(m.hnode, cell), hn = cld(m.hnode_new, cell), V = cld(m.areasvol, cell);
                double sum = 0.0;
                const size_t eo = (size_t)a * Ep;
                for (int i = ep0; i < ep1; ++i) {
                    const int2 en = sEnt[i];
                    if (a + 1 > ent_top(en)) continue;
                    const int e2 = ent_e2(en), oth = ent_oth(en);
                    const bool second = e2 & 1;
                    const D to = sT[nxt][oth], tn = sT[nxt][r];
                    const double f = upw1_edge<D>(second ? to : tn, second ? tn : to, cld(m.q, eo + (e2 >> 1)));
                    sum = second ? sum - f : sum + f;
                }
                lo_new = ((double)tw[2] * h + (sum + ((double)lo_top - (double)lo_bot)) * dt / V) / hn;
                lo_top = lo_bot;
            }
            if (!(act && ring <= 2 && a + 1 <= L)) av[3] = 0;
            lo[2] = lo_new;
            sLO[r] = lo_new;
        }
        __syncthreads();
        // ---- stage 3: rings 0-1: cluster(a+1), adfh(a), R(a)
        if (act && ring <= 1) {
            if (a + 1 <= L) {
                const D l0 = sLO[r], t0 = sT[nxt][r];
                D vmax = cmax(l0, t0), vmin = cmin(l0, t0);
                for (int i = ep0; i < ep1; ++i) {
                    const int2 en = sEnt[i];
                    if (a + 1 > ent_top(en)) continue;
                    const int oth = ent_oth(en);
                    const D lo_o = sLO[oth], tt = sT[nxt][oth];
                    vmax = cmax(vmax, cmax(lo_o, tt));
                    vmin = cmin(vmin, cmin(lo_o, tt));
                }
                if (a + 1 >= dry_from) {
                    vmax = cmax(vmax, D(-1.0e3));
                    vmin = cmin(vmin, D(1.0e3));
                }
                cmx[2] = vmax;
                cmn[2] = vmin;
            } else {
                cmx[2] = cmn[2] = 0;
            }
            if (a >= 1 && a <= L) {
                // antidiffusive edge fluxes of layer a (k_edge), P+/P- (k_bounds)
                const size_t eo = (size_t)(a - 1) * Ep;
                double pp = 0.0 + ((double)cmax(D(0), av[1]) + (double)cmax(D(0), -av[2]));
                double pm = 0.0 + ((double)cmin(D(0), av[1]) + (double)cmin(D(0), -av[2]));
                int k = 0;
                for (int i = ep0; i < ep1; ++i, ++k) {
                    const int2 en = sEnt[i];
                    if (a > ent_top(en)) continue;
                    const int e2 = ent_e2(en), oth = ent_oth(en), e = e2 >> 1;
                    const bool second = e2 & 1;
                    const int l1 = second ? oth : r, l2 = second ? r : oth;
                    const D t1 = sT[cur][l1], t2 = sT[cur][l2];
                    const D v = cld(m.q, eo + e);
                    const D fin = upw1_edge(t1, t2, v);
                    D f;
                    if (HOR == HOR_UPW1) {
                        f = upw1_edge(t1, t2, v) - fin;
                    } else {
                        const int4 q4 = cld(m.edge_ne, e);
                        const int n1 = q4.x, n2 = q4.y;
                        const D dx = cld(m.edxdy, 2 * (size_t)e), dy = cld(m.edxdy, 2 * (size_t)e + 1);
                        D ac = (D)kREarth * cld(m.elem_cos, q4.z);
                        if (q4.w >= 0) ac = D(0.5) * (ac + (D)kREarth * cld(m.elem_cos, q4.w));
                        const int up = ent_up(en), dn = ent_dn(en);
                        int shared_end = 0;
                        if (up != 2047 && dn != 2047) shared_end = min(cld(m.nlev_min, n1), cld(m.nlev_min, n2)) - 1;
                        D gr[4];
                        if (a <= shared_end) {
                            gr[0] = sGx[up];
                            gr[1] = sGx[dn];
                            gr[2] = sGy[up];
                            gr[3] = sGy[dn];
                        } else {
                            tile_mean_grad(m, sEl, sElId, sElp[l1], sElp[l1 + 1], a, sGx, sGy, gr[0], gr[2]);
                            tile_mean_grad(m, sEl, sElId, sElp[l2], sElp[l2 + 1], a, sGx, sGy, gr[1], gr[3]);
                        }
                        D c1 = 1, c2 = 1;
                        if (HOR == HOR_MUSCL) {
                            c1 = (cld(m.nbl, n1) - a >= 0) ? D(1) : D(0);
                            c2 = (cld(m.nbl, n2) - a >= 0) ? D(1) : D(0);
                        }
                        f = ho_edge<HOR, D>(t1, t2, gr, v, dx, dy, ac, (D)num_ord, c1, c2, fin);
                    }
                    if (k < kTileDeg) ah[k] = f;
                    const D sa = second ? -f : f;
                    pp = pp + (double)cmax(D(0), sa);
                    pm = pm + (double)cmin(D(0), sa);
                }
                const size_t cell = (size_t)(a - 1) * Np + n;
                const double V = cld(m.areasvol, cell);
                const D lo_a = lo[1];
                D qmax, qmin;
                if (a == L || a == 1) {
                    qmax = cmx[1] - lo_a;
                    qmin = cmn[1] - lo_a;
                } else {
                    qmax = cmax(cmax(cmx[0], cmx[1]), cmx[2]) - lo_a;
                    qmin = cmin(cmin(cmn[0], cmn[1]), cmn[2]) - lo_a;
                }
                const double fp = pp * dt / V + 1e-16;
                const double fm = pm * dt / V - 1e-16;
                const double rpv = (double)qmax / fp, rmv = (double)qmin / fm;
                rp[2] = rpv < 1.0 ? rpv : 1.0;
                rm[2] = rmv < 1.0 ? rmv : 1.0;
            } else {
                rp[2] = rm[2] = 0;
            }
        }
        sRp[cur][r] = rp[2];
        sRm[cur][r] = rm[2];
        __syncthreads();
        // ---- stage 4: ring 0: limiting, flux divergence and update of layer c = a-1
        const int c = a - 1;
        if (act && ring == 0 && c >= 1 && c <= L) {
            auto limit = [&](D x, int j, D rp0, D rm0, D rp1, D rm1) -> D {
                D ae = 1;
                if (j == 1) ae = x >= D(0) ? cmin(ae, rp1) : cmin(ae, rm1);
                else if (x >= D(0)) {
                    ae = cmin(ae, rm0);
                    ae = cmin(ae, rp1);
                } else {
                    ae = cmin(ae, rp0);
                    ae = cmin(ae, rm1);
                }
                return ae * x;
            };
            const size_t cell = (size_t)(c - 1) * Np + n;
            const double V = cld(m.areasvol, cell), hn = cld(m.hnode_new, cell), h = cld(m.hnode, cell);
            const double tval = tw[0];  // T(a-1)
            // layers c-1, c, c+1 -> rp[0], rp[1], rp[2]
            const D top = limit(av[0], c, rp[0], rm[0], rp[1], rm[1]);
            const D bot = (c + 1 <= L) ? limit(av[1], c + 1, rp[1], rm[1], rp[2], rm[2]) : av[1];
            double dv = 0.0;
            dv = dv - tval * h + (double)lo[0] * hn;
            dv = dv + ((double)top - (double)bot) * dt / V;
            double dh = 0.0;
            const D rpn = rp[1], rmn = rm[1];
            int k = 0;
            for (int i = ep0; i < ep1; ++i, ++k) {
                const int2 en = sEnt[i];
                if (c > ent_top(en)) continue;
                const int e2 = ent_e2(en), oth = ent_oth(en);
                const bool second = e2 & 1;
                D f = k < kTileDeg ? ahp[k] : D(0);
                const D rpo = sRp[c & 1][oth], rmo = sRm[c & 1][oth];
                const D rp_1 = second ? rpo : rpn, rm_1 = second ? rmo : rmn;
                const D rp_2 = second ? rpn : rpo, rm_2 = second ? rmn : rmo;
                D ae = 1;
                if (f >= D(0)) {
                    ae = cmin(ae, rp_1);
                    ae = cmin(ae, rm_2);
                } else {
                    ae = cmin(ae, rm_1);
                    ae = cmin(ae, rp_2);
                }
                f = ae * f;
                dh = second ? dh - (double)f * dt / V : dh + (double)f * dt / V;
            }
            Tout[cell] = tval + (dh + dv) / hn;
        }
        // ---- shift the windows one level down
#pragma unroll
        for (int i = 0; i < kTileDeg; ++i) ahp[i] = ah[i];
        lo[0] = lo[1];
        lo[1] = lo[2];
        cmx[0] = cmx[1];
        cmx[1] = cmx[2];
        cmn[0] = cmn[1];
        cmn[1] = cmn[2];
        rp[0] = rp[1];
        rp[1] = rp[2];
        rm[0] = rm[1];
        rm[1] = rm[2];
        av[0] = av[1];
        av[1] = av[2];
        av[2] = av[3];
    }
}

template <int HOR, int VER>
static void tile_launch(int ntiles, hipStream_t s, const MeshCol<double> &m, const TileMesh &tm, const double *Tin,
                        double *Tout, double ph, double pv, double dt) {
    k_tile<HOR, VER><<<dim3((unsigned)ntiles), kTileThreads, 0, s>>>(m, tm, Tin, Tout, ph, pv, dt);
}

hipError_t launch_tile_step(const MeshCol<double> &m, const TileMesh &tm, const double *Tin, double *Tout, int hor,
                            int ver, double ph, double pv, double dt, hipStream_t s) {
    if (tm.ntiles % kXcds) return hipErrorInvalidValue;  // padded to the XCD count by the host
#define TILE_CASE(H, V)                                                      \
    case H * 4 + V: tile_launch<H, V>(tm.ntiles, s, m, tm, Tin, Tout, ph, pv, dt); break;
    switch (hor * 4 + ver) {
        TILE_CASE(HOR_UPW1, VER_UPW1)
        TILE_CASE(HOR_UPW1, VER_QR4C)
        TILE_CASE(HOR_UPW1, VER_CDIFF)
        TILE_CASE(HOR_UPW1, VER_PPM)
        TILE_CASE(HOR_MUSCL, VER_UPW1)
        TILE_CASE(HOR_MUSCL, VER_QR4C)
        TILE_CASE(HOR_MUSCL, VER_CDIFF)
        TILE_CASE(HOR_MUSCL, VER_PPM)
        TILE_CASE(HOR_MFCT, VER_UPW1)
        TILE_CASE(HOR_MFCT, VER_QR4C)
        TILE_CASE(HOR_MFCT, VER_CDIFF)
        TILE_CASE(HOR_MFCT, VER_PPM)
    default: return hipErrorInvalidValue;
    }
#undef TILE_CASE
    return hipGetLastError();
}

}  // namespace fesomtra
