// Device-side data layout and the per-(entity, level) formulas of the tracer advection step.
//
// Layout in HBM (all 0-based, entities renumbered along a space-filling curve):
//   node-layer arrays   a[n * L + k]        (L = nl-1 layers; k = nz-1)
//   node-interface      a[n * NL + k]       (NL = nl interfaces)
//   element-layer       a[el * L + k], velocities uv[(el * L + k) * 2 + c]
//   edge-layer          a[e * L + k]
// Every formula below keeps the operation order of the reference routine it cites, so the FP64
// kernels (built with -ffp-contract=off) reproduce the serial Fortran dp build bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fesomtra {

constexpr double kREarth = 6367500.0;  // lib/oce_modules.F90:24

enum Hor { HOR_UPW1 = 0, HOR_MUSCL = 1, HOR_MFCT = 2 };
enum Ver { VER_UPW1 = 0, VER_QR4C = 1, VER_CDIFF = 2, VER_PPM = 3 };

struct MeshDev {
    int N, T, E, nl, L, max_adj;
    // nodes
    const int *nlev_n, *ulev_n, *nlev_min, *ulev_max, *nbl;
    const int *ne_ptr;        // CSR node -> edges (N + 1)
    const int2 *ne_list;      // x = edge * 2 + (node is edges(2)), y = (nu12 << 16) | nl12
    const int *nie;           // (N, max_adj) new element ids, reference list order
    const int *nie_num;
    const double *hnode, *hnode_new, *Z;   // (N, L)
    const double *area, *areasvol, *zbar;  // (N, NL)
    // elements
    const int *enodes;        // (T, 3)
    const int *nlev_e, *ulev_e;
    const double *elem_area, *elem_cos, *gsca;  // gsca (T, 6)
    const double *helem;      // (T, L)
    // edges
    const int4 *edge_ne;      // n1, n2, el1, el2 (-1 = none)
    const double4 *ecd;       // edge_cross_dxdy
    const double2 *edxdy;     // edge_dxdy (radians)
    const int2 *updn;         // up / down triangle (-1 = none)
    // dynamics
    const double *uv;         // (T, L, 2)
    const double *w, *wi, *we;  // (N, NL)
};

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; }

// Column tile: a block of blockDim.x threads owns `cols` whole columns of L levels.
struct Tile {
    int col, k, ent;
    bool live;
};
__device__ __forceinline__ Tile tile_of(int L, int cols, int n_ent) {
    Tile t;
    t.col = threadIdx.x / L;
    t.k = threadIdx.x - t.col * L;
    t.ent = blockIdx.x * cols + t.col;
    t.live = (t.col < cols) && (t.ent < n_ent);
    return t;
}

// ------------------------------------------------------------------------------------------
// Horizontal edge fluxes (lib/oce_adv_tra_hor.F90)

struct EdgeGeo {
    int n1, n2, el1, el2, nl1, nu1, nl2, nu2;
    double a;
};

__device__ __forceinline__ EdgeGeo edge_geo(const MeshDev &m, int e) {
    const int4 q = m.edge_ne[e];
    EdgeGeo g;
    g.n1 = q.x;
    g.n2 = q.y;
    g.el1 = q.z;
    g.el2 = q.w;
    g.nl1 = m.nlev_e[g.el1] - 1;
    g.nu1 = m.ulev_e[g.el1];
    g.a = kREarth * m.elem_cos[g.el1];
    g.nl2 = 0;
    g.nu2 = 0;
    if (g.el2 >= 0) {
        g.nl2 = m.nlev_e[g.el2] - 1;
        g.nu2 = m.ulev_e[g.el2];
        g.a = 0.5 * (g.a + kREarth * m.elem_cos[g.el2]);
    }
    return g;
}

// Volume flux through the dual segment(s) active at level nz (segments A-E of
// lib/oce_adv_tra_hor.F90:166-253).
__device__ __forceinline__ double edge_vflux(const MeshDev &m, const EdgeGeo &g, int e, int nz) {
    const int L = m.L, k = nz - 1;
    const int nl12 = min(g.nl1, g.nl2), nu12 = max(g.nu1, g.nu2);
    bool use1, use2;
    if (nz >= g.nu1 && nz <= nu12 - 1) { use1 = true; use2 = false; }                 // (A)
    else if (g.nu2 > 0 && nz >= g.nu2 && nz <= nu12 - 1) { use1 = false; use2 = true; } // (B)
    else if (nz >= nu12 && nz <= nl12) { use1 = true; use2 = true; }                   // (C)
    else if (nz >= nl12 + 1 && nz <= g.nl1) { use1 = true; use2 = false; }             // (D)
    else { use1 = false; use2 = true; }                                                 // (E)
    const double4 cd = m.ecd[e];
    double v1 = 0.0, v2 = 0.0;
    if (use1) {
        const double2 u = *reinterpret_cast<const double2 *>(&m.uv[((size_t)g.el1 * L + k) * 2]);
        v1 = (-u.y * cd.x + u.x * cd.y) * m.helem[(size_t)g.el1 * L + k];
    }
    if (use2) {
        const double2 u = *reinterpret_cast<const double2 *>(&m.uv[((size_t)g.el2 * L + k) * 2]);
        v2 = (u.y * cd.z - u.x * cd.w) * m.helem[(size_t)g.el2 * L + k];
    }
    return (use1 && use2) ? v1 + v2 : (use1 ? v1 : v2);
}

// First-order upwind edge flux, lib/oce_adv_tra_hor.F90:178-181.
__device__ __forceinline__ double upw1_edge(double t1, double t2, double v) {
    return -0.5 * (t1 * (v + fabs(v)) + t2 * (v - fabs(v)));
}

// tr_xy(comp, nz, el): P1 gradient of the nodal field on element el, summed in node order.
__device__ __forceinline__ double tr_xy(const MeshDev &m, const double *__restrict__ t, int el,
                                        int comp, int k) {
    const double *g = m.gsca + (size_t)el * 6 + comp * 3;
    const int *en = m.enodes + (size_t)el * 3;
    const int L = m.L;
    return g[0] * t[(size_t)en[0] * L + k] + g[1] * t[(size_t)en[1] * L + k] +
           g[2] * t[(size_t)en[2] * L + k];
}

// Area-weighted mean element gradient around node n (lib/oce_muscl_adv.F90:394-409).
__device__ __forceinline__ void node_mean_grad(const MeshDev &m, const double *__restrict__ t,
                                               int n, int nz, double &gx, double &gy) {
    double tvol = 0.0, tx = 0.0, ty = 0.0;
    const int k = nz - 1;
    const int cnt = m.nie_num[n];
    for (int i = 0; i < cnt; ++i) {
        const int el = m.nie[(size_t)n * m.max_adj + i];
        if (m.nlev_e[el] - 1 < nz || nz < m.ulev_e[el]) continue;
        const double ar = m.elem_area[el];
        tvol = tvol + ar;
        tx = tx + tr_xy(m, t, el, 0, k) * ar;
        ty = ty + tr_xy(m, t, el, 1, k) * ar;
    }
    gx = tx / tvol;
    gy = ty / tvol;
}

// edge_up_dn_grad(1:4, nz, e) as fill_up_dn_grad leaves it (lib/oce_muscl_adv.F90:383-527).
__device__ __forceinline__ void edge_grad(const MeshDev &m, const double *__restrict__ t,
                                          const EdgeGeo &g, int e, int nz, double gr[4]) {
    const int2 ud = m.updn[e];
    const int k = nz - 1;
    if (ud.x >= 0 && ud.y >= 0) {
        const int nzmin = max(m.ulev_max[g.n1], m.ulev_max[g.n2]);
        const int nzmax = min(m.nlev_min[g.n1], m.nlev_min[g.n2]);
        if (nz >= nzmin && nz <= nzmax - 1) {
            gr[0] = tr_xy(m, t, ud.x, 0, k);
            gr[1] = tr_xy(m, t, ud.y, 0, k);
            gr[2] = tr_xy(m, t, ud.x, 1, k);
            gr[3] = tr_xy(m, t, ud.y, 1, k);
            return;
        }
    }
    node_mean_grad(m, t, g.n1, nz, gr[0], gr[2]);
    node_mean_grad(m, t, g.n2, nz, gr[1], gr[3]);
}

// MUSCL / MFCT high-order flux minus fin (lib/oce_adv_tra_hor.F90:368-382 and 748-790).
template <int HOR>
__device__ __forceinline__ double ho_edge(const MeshDev &m, const EdgeGeo &g, int e, int nz,
                                          double t1, double t2, const double gr[4], double v,
                                          double num_ord, double fin) {
    const double2 d = m.edxdy[e];
    double Tmean1, Tmean2;
    if (HOR == HOR_MUSCL) {
        const double c1 = (m.nbl[g.n1] - nz >= 0) ? 1.0 : 0.0;
        const double c2 = (m.nbl[g.n2] - nz >= 0) ? 1.0 : 0.0;
        Tmean2 = t2 - (2.0 * (t2 - t1) + d.x * g.a * gr[1] + d.y * kREarth * gr[3]) / 6.0 * c2;
        Tmean1 = t1 + (2.0 * (t2 - t1) + d.x * g.a * gr[0] + d.y * kREarth * gr[2]) / 6.0 * c1;
    } else {
        Tmean2 = t2 - (2.0 * (t2 - t1) + d.x * g.a * gr[1] + d.y * kREarth * gr[3]) / 6.0;
        Tmean1 = t1 + (2.0 * (t2 - t1) + d.x * g.a * gr[0] + d.y * kREarth * gr[2]) / 6.0;
    }
    const double cHO = (v + fabs(v)) * Tmean1 + (v - fabs(v)) * Tmean2;
    return -0.5 * (1.0 - num_ord) * cHO - v * num_ord * 0.5 * (Tmean1 + Tmean2) - fin;
}

// Union level range of an edge packed in the CSR entry (lib/oce_adv_tra_driver.F90:125-136).
__device__ __forceinline__ bool in_range(int packed, int nz) {
    return nz >= (packed >> 16) && nz <= (packed & 0xffff);
}

// ------------------------------------------------------------------------------------------
// Vertical column fluxes (lib/oce_adv_tra_ver.F90), 1-based interface index j.

struct Column {
    const double *t;   // layer values, t[nz - 1]
    const double *w;   // interface velocity, w[j - 1]
    const double *ar;  // area(j)
    const double *Z;   // Z_3d_n(nz)
    const double *zb;  // zbar_3d_n(j)
    const double *hn;  // hnode_new(nz)
    const double *ho;  // hnode(nz)
    int nzmin, nzmax;  // ulevels_nod2D, nlevels_nod2D
    __device__ __forceinline__ double T(int nz) const { return t[nz - 1]; }
    __device__ __forceinline__ double W(int j) const { return w[j - 1]; }
    __device__ __forceinline__ double A(int j) const { return ar[j - 1]; }
};

__device__ __forceinline__ Column column(const MeshDev &m, const double *t, const double *w,
                                         int n) {
    Column c;
    c.t = t + (size_t)n * m.L;
    c.w = w + (size_t)n * m.nl;
    c.ar = m.area + (size_t)n * m.nl;
    c.Z = m.Z + (size_t)n * m.L;
    c.zb = m.zbar + (size_t)n * m.nl;
    c.hn = m.hnode_new + (size_t)n * m.L;
    c.ho = m.hnode + (size_t)n * m.L;
    c.nzmin = m.ulev_n[n];
    c.nzmax = m.nlev_n[n];
    return c;
}

// adv_tra_ver_upw1 (:300-327) at interface j, minus fin.
__device__ __forceinline__ double ver_upw1(const Column &c, int j, double fin) {
    if (j == c.nzmin) return -c.W(j) * c.T(j) * c.A(j) - fin;
    if (j == c.nzmax) return 0.0 - fin;
    const double w = c.W(j);
    return -0.5 * (c.T(j) * (w + fabs(w)) + c.T(j - 1) * (w - fabs(w))) * c.A(j) - fin;
}

// adv_tra_ver_qr4c (:394-436) at interface j, minus fin; a 2-layer column updates the middle
// interface twice, exactly as the reference's statement sequence does.
__device__ __forceinline__ double ver_qr4c(const Column &c, int j, double num_ord, double fin) {
    if (j == c.nzmin) return -c.T(j) * c.W(j) * c.A(j) - fin;
    if (j == c.nzmax) return 0.0 - fin;
    const bool second = (j == c.nzmin + 1), penult = (j == c.nzmax - 1);
    if (second || penult) {
        const double cv = -0.5 * (c.T(j - 1) + c.T(j)) * c.W(j) * c.A(j);
        double f = fin;
        if (second) f = cv - f;
        if (penult) f = cv - f;
        return f;
    }
    const double qc = (c.T(j - 1) - c.T(j)) / (c.Z[j - 2] - c.Z[j - 1]);
    const double qu = (c.T(j) - c.T(j + 1)) / (c.Z[j - 1] - c.Z[j]);
    const double qd = (c.T(j - 2) - c.T(j - 1)) / (c.Z[j - 3] - c.Z[j - 2]);
    const double Tmean1 = c.T(j) + (2.0 * qc + qu) * (c.zb[j - 1] - c.Z[j - 1]) / 3.0;
    const double Tmean2 = c.T(j - 1) + (2.0 * qc + qd) * (c.zb[j - 1] - c.Z[j - 2]) / 3.0;
    const double w = c.W(j);
    const double Tmean = (w + fabs(w)) * Tmean1 + (w - fabs(w)) * Tmean2;
    return (-0.5 * (1.0 - num_ord) * Tmean - num_ord * (0.5 * (Tmean1 + Tmean2)) * w) * c.A(j) - fin;
}

// adv_tra_ver_cdiff (:687-708): the bottom interface keeps fin.
__device__ __forceinline__ double ver_cdiff(const Column &c, int j, double fin) {
    if (j == c.nzmax) return fin;
    if (j == c.nzmin) return -c.W(j) * c.T(j) * c.A(j) - fin;
    const double tv = 0.5 * (c.T(j - 1) + c.T(j));
    return -tv * c.W(j) * c.A(j) - fin;
}

// PPM interface value tv(j) (:510-594) after all of the reference's assignments.
__device__ __forceinline__ double ppm_tv(const Column &c, int j) {
    const int nzmin = c.nzmin, nzmax = c.nzmax;
    if (j >= nzmin + 2 && j <= nzmax - 2) {
        const int nz = j - 1;
        const double dzjm1 = c.hn[nz - 2], dzj = c.hn[nz - 1], dzjp1 = c.hn[nz], dzjp2 = c.hn[nz + 1];
        const double tm1 = c.T(nz - 1), t0 = c.T(nz), tp1 = c.T(nz + 1), tp2 = c.T(nz + 2);
        double deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                        ((2.0 * dzjm1 + dzj) / (dzjp1 + dzj) * (tp1 - t0) +
                         (dzj + 2.0 * dzjp1) / (dzjm1 + dzj) * (t0 - tm1));
        double deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                          ((2.0 * dzj + dzjp1) / (dzjp2 + dzjp1) * (tp2 - tp1) +
                           (dzjp1 + 2.0 * dzjp2) / (dzj + dzjp1) * (tp1 - t0));
        if ((tp1 - t0) * (t0 - tm1) > 0.0)
            deltaj = dmin(dmin(fabs(deltaj), 2.0 * fabs(tp1 - t0)), 2.0 * fabs(t0 - tm1)) * fsign1(deltaj);
        else
            deltaj = 0.0;
        if ((tp2 - tp1) * (tp1 - t0) > 0.0)
            deltajp1 = dmin(dmin(fabs(deltajp1), 2.0 * fabs(tp2 - tp1)), 2.0 * fabs(tp1 - t0)) * fsign1(deltajp1);
        else
            deltajp1 = 0.0;
        return t0 + dzj / (dzj + dzjp1) * (tp1 - t0) +
               1.0 / (dzjm1 + dzj + dzjp1 + dzjp2) *
                   ((2.0 * dzjp1 * dzj) / (dzj + dzjp1) *
                        ((dzjm1 + dzj) / (2.0 * dzj + dzjp1) - (dzjp2 + dzjp1) / (2.0 * dzjp1 + dzj)) *
                        (tp1 - t0) -
                    dzj * (dzjm1 + dzj) / (2.0 * dzj + dzjp1) * deltajp1 +
                    dzjp1 * (dzjp1 + dzjp2) / (dzj + 2.0 * dzjp1) * deltaj);
    }
    if (j == nzmax) return c.T(nzmax - 1);
    if (j == nzmax - 1) {
        const double s = fsign1(c.W(nzmax - 1));
        return -c.T(nzmax - 2) * dmin(s, 0.0) + c.T(nzmax - 1) * dmax(s, 0.0);
    }
    if (j == nzmin + 1) return 0.5 * (c.T(nzmin) + c.T(nzmin + 1));
    return c.T(nzmin);
}

// Limited parabola of layer nz (:601-617).
__device__ __forceinline__ void ppm_layer(const Column &c, int nz, double &aL, double &aR,
                                          double &aj) {
    const double t = c.T(nz);
    aL = ppm_tv(c, nz);
    aR = ppm_tv(c, nz + 1);
    if ((aR - t) * (t - aL) <= 0.0) {
        aL = t;
        aR = t;
    }
    if ((aR - aL) * (t - 0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / 6.0) aL = 3.0 * t - 2.0 * aR;
    if ((aR - aL) * (t - 0.5 * (aR + aL)) < -(aR - aL) * (aR - aL) / 6.0) aR = 3.0 * t - 2.0 * aL;
    aj = 6.0 * (t - 0.5 * (aL + aR));
}

// adv_tra_vert_ppm (:596-639) at interface j, minus fin.
__device__ __forceinline__ double ver_ppm(const Column &c, int j, double dt, double fin) {
    if (j == c.nzmin) return -ppm_tv(c, c.nzmin) * c.W(c.nzmin) * c.A(c.nzmin) - fin;
    if (j == c.nzmax) return 0.0 - fin;
    const double wj = c.W(j);
    double tvert = 0.0;
    double aL, aR, aj;
    if (wj > 0.0) {  // top of layer j, upstream below... the reference's "W(nz) > 0" branch
        ppm_layer(c, j, aL, aR, aj);
        const double x = dmin(wj * dt / c.ho[j - 1], 1.0);
        tvert = (-aL - 0.5 * x * (aR - aL + (1.0 - 2.0 / 3.0 * x) * aj));
        tvert = tvert * c.A(j) * wj;
    } else if (wj < 0.0) {  // bottom of layer j-1: the "W(nz+1) < 0" branch
        ppm_layer(c, j - 1, aL, aR, aj);
        const double x = dmin(-wj * dt / c.ho[j - 2], 1.0);
        tvert = (-aR + 0.5 * x * (aR - aL - (1.0 - 2.0 / 3.0 * x) * aj));
        tvert = tvert * c.A(j) * wj;
    }
    return tvert - fin;
}

template <int VER>
__device__ __forceinline__ double ver_ho(const Column &c, int j, double num_ord, double dt,
                                         double fin) {
    if (VER == VER_QR4C) return ver_qr4c(c, j, num_ord, fin);
    if (VER == VER_CDIFF) return ver_cdiff(c, j, fin);
    if (VER == VER_PPM) return ver_ppm(c, j, dt, fin);
    return ver_upw1(c, j, fin);
}

}  // namespace fesomtra
