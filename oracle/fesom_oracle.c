/* CPU restatement of the FESOM2 tracer advection step -- TEST INFRASTRUCTURE ONLY.
 *
 * Follows sbrdar/fesom2_tracer_dwarf line by line (citations per routine); operation order is
 * kept so that, compiled with -ffp-contract=off, the results equal the serial Fortran dp build
 * bit for bit.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Indexing: a Fortran array a(d1, d2[, d3]) with 1-based indices maps to a flat C array through
 * the macros below; all integer index arrays hold 1-based values.
 */
#include "fesom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R_EARTH 6367500.0 /* lib/oce_modules.F90:24 */

#define IX2(d1, i, j) ((size_t)((j) - 1) * (size_t)(d1) + (size_t)((i) - 1))
#define IX3(d1, d2, i, j, k) \
    (((size_t)((k) - 1) * (size_t)(d2) + (size_t)((j) - 1)) * (size_t)(d1) + (size_t)((i) - 1))

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
/* Fortran SIGN(1.0, x) with IEEE signed zero. */
static inline double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; }

/* ------------------------------------------------------------------------------------------ */
/* Element gradients of a nodal field: tr_xy(1:2, nz, elem) over the element's wet levels,
 * sum over its three nodes in node order (the caller's gradient, as FESOM2 computes it from
 * gradient_sca, lib/oce_mesh.F90:2512-2534). */
void fo_tracer_gradients(const fo_mesh *m, const double *ttf, double *tr_xy)
{
    const int L = m->nl - 1;
    for (int el = 1; el <= m->elem2D; ++el) {
        const int n1 = m->elem2D_nodes[IX2(3, 1, el)];
        const int n2 = m->elem2D_nodes[IX2(3, 2, el)];
        const int n3 = m->elem2D_nodes[IX2(3, 3, el)];
        const double *g = &m->gradient_sca[IX2(6, 1, el)];
        for (int nz = m->ulevels[el - 1]; nz <= m->nlevels[el - 1] - 1; ++nz) {
            const double t1 = ttf[IX2(L, nz, n1)], t2 = ttf[IX2(L, nz, n2)], t3 = ttf[IX2(L, nz, n3)];
            tr_xy[IX3(2, L, 1, nz, el)] = g[0] * t1 + g[1] * t2 + g[2] * t3;
            tr_xy[IX3(2, L, 2, nz, el)] = g[3] * t1 + g[4] * t2 + g[5] * t3;
        }
    }
}

/* Area-weighted mean of the element gradients around node n at level nz
 * (lib/oce_muscl_adv.F90:394-409). */
static void node_mean_grad(const fo_mesh *m, const double *tr_xy, int n, int nz, double *gx,
                           double *gy)
{
    const int L = m->nl - 1;
    double tvol = 0.0, tx = 0.0, ty = 0.0;
    for (int k = 1; k <= m->nod_in_elem2D_num[n - 1]; ++k) {
        const int elem = m->nod_in_elem2D[IX2(m->max_adj, k, n)];
        if (m->nlevels[elem - 1] - 1 < nz || nz < m->ulevels[elem - 1]) continue;
        tvol = tvol + m->elem_area[elem - 1];
        tx = tx + tr_xy[IX3(2, L, 1, nz, elem)] * m->elem_area[elem - 1];
        ty = ty + tr_xy[IX3(2, L, 2, nz, elem)] * m->elem_area[elem - 1];
    }
    *gx = tx / tvol;
    *gy = ty / tvol;
}

/* lib/oce_muscl_adv.F90:361-530 */
void fo_fill_up_dn_grad(const fo_mesh *m, const double *tr_xy, double *g)
{
    const int L = m->nl - 1;
    double gx, gy;
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        const int n1 = m->edges[IX2(2, 1, edge)], n2 = m->edges[IX2(2, 2, edge)];
        const int up = m->edge_up_dn_tri[IX2(2, 1, edge)], dn = m->edge_up_dn_tri[IX2(2, 2, edge)];
        if (up != 0 && dn != 0) {
            const int nzmin = imax(m->ulevels_nod2D_max[n1 - 1], m->ulevels_nod2D_max[n2 - 1]);
            const int nzmax = imin(m->nlevels_nod2D_min[n1 - 1], m->nlevels_nod2D_min[n2 - 1]);
            for (int nz = m->ulevels_nod2D[n1 - 1]; nz <= nzmin - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                g[IX3(4, L, 1, nz, edge)] = gx;
                g[IX3(4, L, 3, nz, edge)] = gy;
            }
            for (int nz = m->ulevels_nod2D[n2 - 1]; nz <= nzmin - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                g[IX3(4, L, 2, nz, edge)] = gx;
                g[IX3(4, L, 4, nz, edge)] = gy;
            }
            for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
                g[IX3(4, L, 1, nz, edge)] = tr_xy[IX3(2, L, 1, nz, up)];
                g[IX3(4, L, 2, nz, edge)] = tr_xy[IX3(2, L, 1, nz, dn)];
                g[IX3(4, L, 3, nz, edge)] = tr_xy[IX3(2, L, 2, nz, up)];
                g[IX3(4, L, 4, nz, edge)] = tr_xy[IX3(2, L, 2, nz, dn)];
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n1 - 1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                g[IX3(4, L, 1, nz, edge)] = gx;
                g[IX3(4, L, 3, nz, edge)] = gy;
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n2 - 1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                g[IX3(4, L, 2, nz, edge)] = gx;
                g[IX3(4, L, 4, nz, edge)] = gy;
            }
        } else {
            for (int nz = m->ulevels_nod2D[n1 - 1]; nz <= m->nlevels_nod2D[n1 - 1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                g[IX3(4, L, 1, nz, edge)] = gx;
                g[IX3(4, L, 3, nz, edge)] = gy;
            }
            for (int nz = m->ulevels_nod2D[n2 - 1]; nz <= m->nlevels_nod2D[n2 - 1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                g[IX3(4, L, 2, nz, edge)] = gx;
                g[IX3(4, L, 4, nz, edge)] = gy;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Edge geometry common to the three horizontal schemes (lib/oce_adv_tra_hor.F90:123-164). */
typedef struct {
    int n1, n2, el1, el2, nl1, nu1, nl2, nu2, nl12, nu12;
    double dX1, dY1, dX2, dY2, a;
} edge_geo;

static void edge_setup(const fo_mesh *m, int edge, edge_geo *g)
{
    g->n1 = m->edges[IX2(2, 1, edge)];
    g->n2 = m->edges[IX2(2, 2, edge)];
    g->el1 = m->edge_tri[IX2(2, 1, edge)];
    g->el2 = m->edge_tri[IX2(2, 2, edge)];
    g->nl1 = m->nlevels[g->el1 - 1] - 1;
    g->nu1 = m->ulevels[g->el1 - 1];
    g->dX1 = m->edge_cross_dxdy[IX2(4, 1, edge)];
    g->dY1 = m->edge_cross_dxdy[IX2(4, 2, edge)];
    g->a = R_EARTH * m->elem_cos[g->el1 - 1];
    g->nl2 = 0;
    g->nu2 = 0;
    g->dX2 = 0.0;
    g->dY2 = 0.0;
    if (g->el2 > 0) {
        g->dX2 = m->edge_cross_dxdy[IX2(4, 3, edge)];
        g->dY2 = m->edge_cross_dxdy[IX2(4, 4, edge)];
        g->nl2 = m->nlevels[g->el2 - 1] - 1;
        g->nu2 = m->ulevels[g->el2 - 1];
        g->a = 0.5 * (g->a + R_EARTH * m->elem_cos[g->el2 - 1]);
    }
    g->nl12 = imin(g->nl1, g->nl2);
    g->nu12 = imax(g->nu1, g->nu2);
}

/* Volume flux through the dual segment of element side 1 or 2 (lib/oce_adv_tra_hor.F90:174,194). */
static inline double vflux1(const fo_mesh *m, const double *vel, const edge_geo *g, int nz)
{
    const int L = m->nl - 1;
    return (-vel[IX3(2, L, 2, nz, g->el1)] * g->dX1 + vel[IX3(2, L, 1, nz, g->el1)] * g->dY1) *
           m->helem[IX2(L, nz, g->el1)];
}
static inline double vflux2(const fo_mesh *m, const double *vel, const edge_geo *g, int nz)
{
    const int L = m->nl - 1;
    return (vel[IX3(2, L, 2, nz, g->el2)] * g->dX2 - vel[IX3(2, L, 1, nz, g->el2)] * g->dY2) *
           m->helem[IX2(L, nz, g->el2)];
}

static inline double upw1_flux(double t1, double t2, double vflux, double fin)
{
    return -0.5 * (t1 * (vflux + fabs(vflux)) + t2 * (vflux - fabs(vflux))) - fin;
}

/* lib/oce_adv_tra_hor.F90:64-261 */
void fo_adv_tra_hor_upw1(const fo_mesh *m, const double *vel, const double *ttf, double *flux,
                         int init_zero)
{
    const int L = m->nl - 1;
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)L * (size_t)m->edge2D);
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        edge_geo g;
        edge_setup(m, edge, &g);
        double *f = &flux[IX2(L, 1, edge)] - 1; /* f[nz], 1-based */
        const double *t1 = &ttf[IX2(L, 1, g.n1)] - 1, *t2 = &ttf[IX2(L, 1, g.n2)] - 1;
        for (int nz = g.nu1; nz <= g.nu12 - 1; ++nz)                      /* (A) */
            f[nz] = upw1_flux(t1[nz], t2[nz], vflux1(m, vel, &g, nz), f[nz]);
        if (g.nu2 > 0)
            for (int nz = g.nu2; nz <= g.nu12 - 1; ++nz)                  /* (B) */
                f[nz] = upw1_flux(t1[nz], t2[nz], vflux2(m, vel, &g, nz), f[nz]);
        for (int nz = g.nu12; nz <= g.nl12; ++nz)                         /* (C) */
            f[nz] = upw1_flux(t1[nz], t2[nz], vflux1(m, vel, &g, nz) + vflux2(m, vel, &g, nz), f[nz]);
        for (int nz = g.nl12 + 1; nz <= g.nl1; ++nz)                      /* (D) */
            f[nz] = upw1_flux(t1[nz], t2[nz], vflux1(m, vel, &g, nz), f[nz]);
        for (int nz = g.nl12 + 1; nz <= g.nl2; ++nz)                      /* (E) */
            f[nz] = upw1_flux(t1[nz], t2[nz], vflux2(m, vel, &g, nz), f[nz]);
    }
}

/* High-order edge flux with the MUSCL/MFCT reconstruction (lib/oce_adv_tra_hor.F90:368-382,
 * 748-790); c1, c2 are MUSCL's c_lo switches (1 for MFCT). */
static inline double ho_flux(const fo_mesh *m, const edge_geo *g, int edge, double t1, double t2,
                             const double *gr, double vflux, double num_ord, double c1,
                             double c2, double fin, int muscl)
{
    const double dx = m->edge_dxdy[IX2(2, 1, edge)], dy = m->edge_dxdy[IX2(2, 2, edge)];
    double Tmean2, Tmean1;
    if (muscl) {
        Tmean2 = t2 - (2.0 * (t2 - t1) + dx * g->a * gr[1] + dy * R_EARTH * gr[3]) / 6.0 * c2;
        Tmean1 = t1 + (2.0 * (t2 - t1) + dx * g->a * gr[0] + dy * R_EARTH * gr[2]) / 6.0 * c1;
    } else {
        Tmean2 = t2 - (2.0 * (t2 - t1) + dx * g->a * gr[1] + dy * R_EARTH * gr[3]) / 6.0;
        Tmean1 = t1 + (2.0 * (t2 - t1) + dx * g->a * gr[0] + dy * R_EARTH * gr[2]) / 6.0;
    }
    const double cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
    return -0.5 * (1.0 - num_ord) * cHO - vflux * num_ord * 0.5 * (Tmean1 + Tmean2) - fin;
}

static void hor_ho(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                   double *flux, const double *egrad, int init_zero, int muscl)
{
    const int L = m->nl - 1;
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)L * (size_t)m->edge2D);
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        edge_geo g;
        edge_setup(m, edge, &g);
        double *f = &flux[IX2(L, 1, edge)] - 1;
        const double *t1 = &ttf[IX2(L, 1, g.n1)] - 1, *t2 = &ttf[IX2(L, 1, g.n2)] - 1;
        const int nb1 = m->nboundary_lay[g.n1 - 1], nb2 = m->nboundary_lay[g.n2 - 1];
#define CLO(nb, nz) ((nb) - (nz) >= 0 ? 1.0 : 0.0)
#define GR(nz) (&egrad[IX3(4, L, 1, (nz), edge)])
        for (int nz = g.nu1; nz <= g.nu12 - 1; ++nz)
            f[nz] = ho_flux(m, &g, edge, t1[nz], t2[nz], GR(nz), vflux1(m, vel, &g, nz), num_ord,
                            CLO(nb1, nz), CLO(nb2, nz), f[nz], muscl);
        if (g.nu2 > 0)
            for (int nz = g.nu2; nz <= g.nu12 - 1; ++nz)
                f[nz] = ho_flux(m, &g, edge, t1[nz], t2[nz], GR(nz), vflux2(m, vel, &g, nz),
                                num_ord, CLO(nb1, nz), CLO(nb2, nz), f[nz], muscl);
        for (int nz = g.nu12; nz <= g.nl12; ++nz)
            f[nz] = ho_flux(m, &g, edge, t1[nz], t2[nz], GR(nz),
                            vflux1(m, vel, &g, nz) + vflux2(m, vel, &g, nz), num_ord,
                            CLO(nb1, nz), CLO(nb2, nz), f[nz], muscl);
        for (int nz = g.nl12 + 1; nz <= g.nl1; ++nz)
            f[nz] = ho_flux(m, &g, edge, t1[nz], t2[nz], GR(nz), vflux1(m, vel, &g, nz), num_ord,
                            CLO(nb1, nz), CLO(nb2, nz), f[nz], muscl);
        for (int nz = g.nl12 + 1; nz <= g.nl2; ++nz)
            f[nz] = ho_flux(m, &g, edge, t1[nz], t2[nz], GR(nz), vflux2(m, vel, &g, nz), num_ord,
                            CLO(nb1, nz), CLO(nb2, nz), f[nz], muscl);
#undef CLO
#undef GR
    }
}

/* lib/oce_adv_tra_hor.F90:265-550 */
void fo_adv_tra_hor_muscl(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                          double *flux, const double *egrad, int init_zero)
{
    hor_ho(m, vel, ttf, num_ord, flux, egrad, init_zero, 1);
}

/* lib/oce_adv_tra_hor.F90:554-846 */
void fo_adv_tra_hor_mfct(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                         double *flux, const double *egrad, int init_zero)
{
    hor_ho(m, vel, ttf, num_ord, flux, egrad, init_zero, 0);
}

/* ------------------------------------------------------------------------------------------ */
/* lib/oce_adv_tra_ver.F90:90-244: implicit first-order upwind vertical advection of ttf (the
 * low-order solution) with the implicit part w of the split vertical velocity, one tridiagonal
 * solve per column on hnode_new (the reference also builds zbar_n / Z_n there, which the solve
 * never reads).  The reference's single-layer branch reads ttf outside the column, so columns
 * with fewer than two wet layers are rejected (-1). */
int fo_adv_tra_vert_impl(const fo_mesh *m, double dt, const double *W, double *ttf)
{
    const int L = m->nl - 1, NL = m->nl;
    double *a = (double *)calloc((size_t)NL + 1, sizeof(double));
    double *b = (double *)calloc((size_t)NL + 1, sizeof(double));
    double *c = (double *)calloc((size_t)NL + 1, sizeof(double));
    double *tr = (double *)calloc((size_t)NL + 1, sizeof(double));
    double *cp = (double *)calloc((size_t)NL + 1, sizeof(double));
    double *tp = (double *)calloc((size_t)NL + 1, sizeof(double));
    int rc = (a && b && c && tr && cp && tp) ? 0 : -1;
    for (int n = 1; n <= m->nod2D && rc == 0; ++n) {
        const int nzmax = m->nlevels_nod2D[n - 1], nzmin = m->ulevels_nod2D[n - 1];
        if (nzmax - 1 - nzmin < 1) {
            rc = -1;
            break;
        }
        const double *w = &W[IX2(NL, 1, n)] - 1, *ar = &m->area[IX2(NL, 1, n)] - 1;
        const double *av = &m->areasvol[IX2(NL, 1, n)] - 1, *hn = &m->hnode_new[IX2(L, 1, n)] - 1;
        double *t = &ttf[IX2(L, 1, n)] - 1;
        for (int k = 0; k <= NL; ++k) a[k] = b[k] = c[k] = tr[k] = cp[k] = tp[k] = 0.0;
        const double zinv = 1.0 * dt;
        double v_adv, dz, mm;
        int nz = nzmin;
        a[nz] = 0.0;
        v_adv = zinv * ar[nz] / av[nz];
        b[nz] = hn[nz] + w[nz] * v_adv;
        v_adv = zinv * ar[nz + 1] / av[nz];
        b[nz] = b[nz] - dmin(0.0, w[nz + 1]) * v_adv;
        c[nz] = -dmax(0.0, w[nz + 1]) * v_adv;
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            v_adv = zinv * ar[nz] / av[nz];
            a[nz] = dmin(0.0, w[nz]) * v_adv;
            b[nz] = hn[nz] + dmax(0.0, w[nz]) * v_adv;
            v_adv = zinv * ar[nz + 1] / av[nz];
            b[nz] = b[nz] - dmin(0.0, w[nz + 1]) * v_adv;
            c[nz] = -dmax(0.0, w[nz + 1]) * v_adv;
        }
        nz = nzmax - 1;
        v_adv = zinv * ar[nz] / av[nz];
        a[nz] = dmin(0.0, w[nz]) * v_adv;
        b[nz] = hn[nz] + dmax(0.0, w[nz]) * v_adv;
        c[nz] = 0.0;
        nz = nzmin;
        dz = hn[nz];
        tr[nz] = -(b[nz] - dz) * t[nz] - c[nz] * t[nz + 1];
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            dz = hn[nz];
            tr[nz] = -a[nz] * t[nz - 1] - (b[nz] - dz) * t[nz] - c[nz] * t[nz + 1];
        }
        nz = nzmax - 1;
        dz = hn[nz];
        tr[nz] = -a[nz] * t[nz - 1] - (b[nz] - dz) * t[nz];
        nz = nzmin;
        cp[nz] = c[nz] / b[nz];
        tp[nz] = tr[nz] / b[nz];
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            mm = b[nz] - cp[nz - 1] * a[nz];
            cp[nz] = c[nz] / mm;
            tp[nz] = (tr[nz] - tp[nz - 1] * a[nz]) / mm;
        }
        tr[nzmax - 1] = tp[nzmax - 1];
        for (nz = nzmax - 2; nz >= nzmin; --nz) tr[nz] = tp[nz] - cp[nz] * tr[nz + 1];
        for (nz = nzmin; nz <= nzmax - 1; ++nz) t[nz] = t[nz] + tr[nz];
    }
    free(a); free(b); free(c); free(tr); free(cp); free(tp);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* lib/oce_adv_tra_ver.F90:248-335 */
void fo_adv_tra_ver_upw1(const fo_mesh *m, const double *W, const double *ttf, double *flux,
                         int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)NL * (size_t)m->nod2D);
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n - 1], nzmin = m->ulevels_nod2D[n - 1];
        double *f = &flux[IX2(NL, 1, n)] - 1;
        const double *w = &W[IX2(NL, 1, n)] - 1, *ar = &m->area[IX2(NL, 1, n)] - 1;
        const double *t = &ttf[IX2(L, 1, n)] - 1;
        int nz = nzmin;
        f[nz] = -w[nz] * t[nz] * ar[nz] - f[nz];
        nz = nzmax;
        f[nz] = 0.0 - f[nz];
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz)
            f[nz] = -0.5 * (t[nz] * (w[nz] + fabs(w[nz])) + t[nz - 1] * (w[nz] - fabs(w[nz]))) * ar[nz] - f[nz];
    }
}

/* lib/oce_adv_tra_ver.F90:339-444 */
void fo_adv_tra_ver_qr4c(const fo_mesh *m, const double *W, const double *ttf, double num_ord,
                         double *flux, int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)NL * (size_t)m->nod2D);
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n - 1], nzmin = m->ulevels_nod2D[n - 1];
        double *f = &flux[IX2(NL, 1, n)] - 1;
        const double *w = &W[IX2(NL, 1, n)] - 1, *ar = &m->area[IX2(NL, 1, n)] - 1;
        const double *t = &ttf[IX2(L, 1, n)] - 1;
        const double *Z = &m->Z_3d_n[IX2(L, 1, n)] - 1, *zb = &m->zbar_3d_n[IX2(NL, 1, n)] - 1;
        int nz = nzmin;
        f[nz] = -t[nz] * w[nz] * ar[nz] - f[nz];
        nz = nzmin + 1;
        f[nz] = -0.5 * (t[nz - 1] + t[nz]) * w[nz] * ar[nz] - f[nz];
        nz = nzmax - 1;
        f[nz] = -0.5 * (t[nz - 1] + t[nz]) * w[nz] * ar[nz] - f[nz];
        nz = nzmax;
        f[nz] = 0.0 - f[nz];
        for (nz = nzmin + 2; nz <= nzmax - 2; ++nz) {
            const double qc = (t[nz - 1] - t[nz]) / (Z[nz - 1] - Z[nz]);
            const double qu = (t[nz] - t[nz + 1]) / (Z[nz] - Z[nz + 1]);
            const double qd = (t[nz - 2] - t[nz - 1]) / (Z[nz - 2] - Z[nz - 1]);
            const double Tmean1 = t[nz] + (2.0 * qc + qu) * (zb[nz] - Z[nz]) / 3.0;
            const double Tmean2 = t[nz - 1] + (2.0 * qc + qd) * (zb[nz] - Z[nz - 1]) / 3.0;
            const double Tmean = (w[nz] + fabs(w[nz])) * Tmean1 + (w[nz] - fabs(w[nz])) * Tmean2;
            f[nz] = (-0.5 * (1.0 - num_ord) * Tmean - num_ord * (0.5 * (Tmean1 + Tmean2)) * w[nz]) * ar[nz] - f[nz];
        }
    }
}

/* lib/oce_adv_tra_ver.F90:448-644 */
void fo_adv_tra_vert_ppm(const fo_mesh *m, double dt, const double *W, const double *ttf,
                         double *flux, int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(NL + 1));
    double *tv = (double *)malloc(sizeof(double) * (size_t)(NL + 1));
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)NL * (size_t)m->nod2D);
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n - 1], nzmin = m->ulevels_nod2D[n - 1];
        double *f = &flux[IX2(NL, 1, n)] - 1;
        const double *w = &W[IX2(NL, 1, n)] - 1, *ar = &m->area[IX2(NL, 1, n)] - 1;
        const double *t = &ttf[IX2(L, 1, n)] - 1;
        const double *hn = &m->hnode_new[IX2(L, 1, n)] - 1, *ho = &m->hnode[IX2(L, 1, n)] - 1;
        tv[nzmin] = t[nzmin];
        tv[nzmin + 1] = 0.5 * (t[nzmin] + t[nzmin + 1]);
        tv[nzmax - 1] = -t[nzmax - 2] * dmin(fsign1(w[nzmax - 1]), 0.0) + t[nzmax - 1] * dmax(fsign1(w[nzmax - 1]), 0.0);
        tv[nzmax] = t[nzmax - 1];
        for (int nz = nzmin + 1; nz <= nzmax - 3; ++nz) {
            const double dzjm1 = hn[nz - 1], dzj = hn[nz], dzjp1 = hn[nz + 1], dzjp2 = hn[nz + 2];
            double deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                            ((2.0 * dzjm1 + dzj) / (dzjp1 + dzj) * (t[nz + 1] - t[nz]) +
                             (dzj + 2.0 * dzjp1) / (dzjm1 + dzj) * (t[nz] - t[nz - 1]));
            double deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                              ((2.0 * dzj + dzjp1) / (dzjp2 + dzjp1) * (t[nz + 2] - t[nz + 1]) +
                               (dzjp1 + 2.0 * dzjp2) / (dzj + dzjp1) * (t[nz + 1] - t[nz]));
            if ((t[nz + 1] - t[nz]) * (t[nz] - t[nz - 1]) > 0.0)
                deltaj = dmin(dmin(fabs(deltaj), 2.0 * fabs(t[nz + 1] - t[nz])), 2.0 * fabs(t[nz] - t[nz - 1])) * fsign1(deltaj);
            else
                deltaj = 0.0;
            if ((t[nz + 2] - t[nz + 1]) * (t[nz + 1] - t[nz]) > 0.0)
                deltajp1 = dmin(dmin(fabs(deltajp1), 2.0 * fabs(t[nz + 2] - t[nz + 1])), 2.0 * fabs(t[nz + 1] - t[nz])) * fsign1(deltajp1);
            else
                deltajp1 = 0.0;
            tv[nz + 1] = t[nz] + dzj / (dzj + dzjp1) * (t[nz + 1] - t[nz]) +
                         1.0 / (dzjm1 + dzj + dzjp1 + dzjp2) *
                             ((2.0 * dzjp1 * dzj) / (dzj + dzjp1) *
                                  ((dzjm1 + dzj) / (2.0 * dzj + dzjp1) - (dzjp2 + dzjp1) / (2.0 * dzjp1 + dzj)) *
                                  (t[nz + 1] - t[nz]) -
                              dzj * (dzjm1 + dzj) / (2.0 * dzj + dzjp1) * deltajp1 +
                              dzjp1 * (dzjp1 + dzjp2) / (dzj + 2.0 * dzjp1) * deltaj);
        }
        for (int k = 1; k <= nzmax; ++k) tvert[k] = 0.0;
        for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
            if (w[nz] <= 0.0 && w[nz + 1] >= 0.0) continue;
            double aL = tv[nz], aR = tv[nz + 1];
            if ((aR - t[nz]) * (t[nz] - aL) <= 0.0) {
                aL = t[nz];
                aR = t[nz];
            }
            if ((aR - aL) * (t[nz] - 0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / 6.0) aL = 3.0 * t[nz] - 2.0 * aR;
            if ((aR - aL) * (t[nz] - 0.5 * (aR + aL)) < -(aR - aL) * (aR - aL) / 6.0) aR = 3.0 * t[nz] - 2.0 * aL;
            const double dzj = ho[nz];
            const double aj = 6.0 * (t[nz] - 0.5 * (aL + aR));
            if (w[nz] > 0.0) {
                const double x = dmin(w[nz] * dt / dzj, 1.0);
                tvert[nz] = (-aL - 0.5 * x * (aR - aL + (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz] = tvert[nz] * ar[nz] * w[nz];
            }
            if (w[nz + 1] < 0.0) {
                const double x = dmin(-w[nz + 1] * dt / dzj, 1.0);
                tvert[nz + 1] = (-aR + 0.5 * x * (aR - aL - (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz + 1] = tvert[nz + 1] * ar[nz + 1] * w[nz + 1];
            }
        }
        tvert[nzmin] = -tv[nzmin] * w[nzmin] * ar[nzmin];
        tvert[nzmax] = 0.0;
        for (int k = nzmin; k <= nzmax; ++k) f[k] = tvert[k] - f[k];
    }
    free(tvert);
    free(tv);
}

/* lib/oce_adv_tra_ver.F90:648-712 */
void fo_adv_tra_ver_cdiff(const fo_mesh *m, const double *W, const double *ttf, double *flux,
                          int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(NL + 2));
    if (init_zero) memset(flux, 0, sizeof(double) * (size_t)NL * (size_t)m->nod2D);
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n - 1] - 1, nzmin = m->ulevels_nod2D[n - 1];
        double *f = &flux[IX2(NL, 1, n)] - 1;
        const double *w = &W[IX2(NL, 1, n)] - 1, *ar = &m->area[IX2(NL, 1, n)] - 1;
        const double *t = &ttf[IX2(L, 1, n)] - 1;
        tvert[nzmin] = -w[nzmin] * t[nzmin] * ar[nzmin];
        tvert[nzmax + 1] = 0.0;
        for (int nz = nzmin + 1; nz <= nzmax; ++nz) {
            const double tv = 0.5 * (t[nz - 1] + t[nz]);
            tvert[nz] = -tv * w[nz] * ar[nz];
        }
        for (int k = nzmin; k <= nzmax; ++k) f[k] = tvert[k] - f[k];
    }
    free(tvert);
}

/* ------------------------------------------------------------------------------------------ */
/* Union level range of an edge used by the scatters (lib/oce_adv_tra_driver.F90:125-136). */
static inline void edge_union_range(const fo_mesh *m, int edge, int *nu12, int *nl12)
{
    const int el1 = m->edge_tri[IX2(2, 1, edge)], el2 = m->edge_tri[IX2(2, 2, edge)];
    const int nl1 = m->nlevels[el1 - 1] - 1, nu1 = m->ulevels[el1 - 1];
    int nl2 = 0, nu2 = 0;
    if (el2 > 0) {
        nl2 = m->nlevels[el2 - 1] - 1;
        nu2 = m->ulevels[el2 - 1];
    }
    *nl12 = imax(nl1, nl2);
    *nu12 = nu1;
    if (nu2 > 0) *nu12 = imin(nu1, nu2);
}

/* lib/oce_adv_tra_fct.F90:53-483 */
void fo_oce_tra_adv_fct(const fo_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, fo_work *wk)
{
    const int L = m->nl - 1, NL = m->nl;
    const double flux_eps = 1e-16, bignumber = 1e3;
    double *fmax = wk->fct_ttf_max, *fmin = wk->fct_ttf_min;
    double *fplus = wk->fct_plus, *fminus = wk->fct_minus;
    double *AUX = wk->edge_up_dn_grad;
    double *tmax = wk->tvert_max, *tmin = wk->tvert_min;

    /* a1 */
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz) {
            fmax[IX2(L, nz, n)] = dmax(lo[IX2(L, nz, n)], ttf[IX2(L, nz, n)]);
            fmin[IX2(L, nz, n)] = dmin(lo[IX2(L, nz, n)], ttf[IX2(L, nz, n)]);
        }
    /* a2 */
    for (int elem = 1; elem <= m->elem2D; ++elem) {
        const int e1 = m->elem2D_nodes[IX2(3, 1, elem)], e2 = m->elem2D_nodes[IX2(3, 2, elem)],
                  e3 = m->elem2D_nodes[IX2(3, 3, elem)];
        const int nu1 = m->ulevels[elem - 1], nl1 = m->nlevels[elem - 1];
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            AUX[IX3(4, L, 1, nz, elem)] = dmax(dmax(fmax[IX2(L, nz, e1)], fmax[IX2(L, nz, e2)]), fmax[IX2(L, nz, e3)]);
            AUX[IX3(4, L, 2, nz, elem)] = dmin(dmin(fmin[IX2(L, nz, e1)], fmin[IX2(L, nz, e2)]), fmin[IX2(L, nz, e3)]);
        }
        if (nl1 <= m->nl - 1)
            for (int nz = nl1; nz <= m->nl - 1; ++nz) {
                AUX[IX3(4, L, 1, nz, elem)] = -bignumber;
                AUX[IX3(4, L, 2, nz, elem)] = bignumber;
            }
    }
    /* a3: cluster bounds */
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz) {
            const int first = m->nod_in_elem2D[IX2(m->max_adj, 1, n)];
            double vmax = AUX[IX3(4, L, 1, nz, first)], vmin = AUX[IX3(4, L, 2, nz, first)];
            for (int k = 2; k <= m->nod_in_elem2D_num[n - 1]; ++k) {
                const int el = m->nod_in_elem2D[IX2(m->max_adj, k, n)];
                vmax = dmax(vmax, AUX[IX3(4, L, 1, nz, el)]);
                vmin = dmin(vmin, AUX[IX3(4, L, 2, nz, el)]);
            }
            tmax[IX2(L, nz, n)] = vmax;
            tmin[IX2(L, nz, n)] = vmin;
        }
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nu1 = m->ulevels_nod2D[n - 1], nl1 = m->nlevels_nod2D[n - 1];
        fmax[IX2(L, nu1, n)] = tmax[IX2(L, nu1, n)] - lo[IX2(L, nu1, n)];
        fmin[IX2(L, nu1, n)] = tmin[IX2(L, nu1, n)] - lo[IX2(L, nu1, n)];
        for (int nz = nu1 + 1; nz <= nl1 - 2; ++nz) {
            fmax[IX2(L, nz, n)] = dmax(dmax(tmax[IX2(L, nz - 1, n)], tmax[IX2(L, nz, n)]), tmax[IX2(L, nz + 1, n)]) - lo[IX2(L, nz, n)];
            fmin[IX2(L, nz, n)] = dmin(dmin(tmin[IX2(L, nz - 1, n)], tmin[IX2(L, nz, n)]), tmin[IX2(L, nz + 1, n)]) - lo[IX2(L, nz, n)];
        }
        const int nz = nl1 - 1;
        fmax[IX2(L, nz, n)] = tmax[IX2(L, nz, n)] - lo[IX2(L, nz, n)];
        fmin[IX2(L, nz, n)] = tmin[IX2(L, nz, n)] - lo[IX2(L, nz, n)];
    }
    /* b1 */
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz) {
            fplus[IX2(L, nz, n)] = 0.0;
            fminus[IX2(L, nz, n)] = 0.0;
        }
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz) {
            const double a = adf_v[IX2(NL, nz, n)], b = adf_v[IX2(NL, nz + 1, n)];
            fplus[IX2(L, nz, n)] = fplus[IX2(L, nz, n)] + (dmax(0.0, a) + dmax(0.0, -b));
            fminus[IX2(L, nz, n)] = fminus[IX2(L, nz, n)] + (dmin(0.0, a) + dmin(0.0, -b));
        }
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        const int n1 = m->edges[IX2(2, 1, edge)], n2 = m->edges[IX2(2, 2, edge)];
        int nu12, nl12;
        edge_union_range(m, edge, &nu12, &nl12);
        for (int nz = nu12; nz <= nl12; ++nz) {
            const double a = adf_h[IX2(L, nz, edge)];
            fplus[IX2(L, nz, n1)] = fplus[IX2(L, nz, n1)] + dmax(0.0, a);
            fminus[IX2(L, nz, n1)] = fminus[IX2(L, nz, n1)] + dmin(0.0, a);
            fplus[IX2(L, nz, n2)] = fplus[IX2(L, nz, n2)] + dmax(0.0, -a);
            fminus[IX2(L, nz, n2)] = fminus[IX2(L, nz, n2)] + dmin(0.0, -a);
        }
    }
    /* b2 */
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz) {
            double flux = fplus[IX2(L, nz, n)] * dt / m->areasvol[IX2(NL, nz, n)] + flux_eps;
            fplus[IX2(L, nz, n)] = dmin(1.0, fmax[IX2(L, nz, n)] / flux);
            flux = fminus[IX2(L, nz, n)] * dt / m->areasvol[IX2(NL, nz, n)] - flux_eps;
            fminus[IX2(L, nz, n)] = dmin(1.0, fmin[IX2(L, nz, n)] / flux);
        }
    /* b3: vertical */
    for (int n = 1; n <= m->nod2D; ++n) {
        const int nu1 = m->ulevels_nod2D[n - 1], nl1 = m->nlevels_nod2D[n - 1];
        int nz = nu1;
        double ae = 1.0;
        double flux = adf_v[IX2(NL, nz, n)];
        if (flux >= 0.0) ae = dmin(ae, fplus[IX2(L, nz, n)]);
        else ae = dmin(ae, fminus[IX2(L, nz, n)]);
        adf_v[IX2(NL, nz, n)] = ae * adf_v[IX2(NL, nz, n)];
        for (nz = nu1 + 1; nz <= nl1 - 1; ++nz) {
            ae = 1.0;
            flux = adf_v[IX2(NL, nz, n)];
            if (flux >= 0.0) {
                ae = dmin(ae, fminus[IX2(L, nz - 1, n)]);
                ae = dmin(ae, fplus[IX2(L, nz, n)]);
            } else {
                ae = dmin(ae, fplus[IX2(L, nz - 1, n)]);
                ae = dmin(ae, fminus[IX2(L, nz, n)]);
            }
            adf_v[IX2(NL, nz, n)] = ae * adf_v[IX2(NL, nz, n)];
        }
    }
    /* b3: horizontal */
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        const int n1 = m->edges[IX2(2, 1, edge)], n2 = m->edges[IX2(2, 2, edge)];
        int nu12, nl12;
        edge_union_range(m, edge, &nu12, &nl12);
        for (int nz = nu12; nz <= nl12; ++nz) {
            double ae = 1.0;
            const double flux = adf_h[IX2(L, nz, edge)];
            if (flux >= 0.0) {
                ae = dmin(ae, fplus[IX2(L, nz, n1)]);
                ae = dmin(ae, fminus[IX2(L, nz, n2)]);
            } else {
                ae = dmin(ae, fminus[IX2(L, nz, n1)]);
                ae = dmin(ae, fplus[IX2(L, nz, n2)]);
            }
            adf_h[IX2(L, nz, edge)] = ae * adf_h[IX2(L, nz, edge)];
        }
    }
}

/* lib/oce_adv_tra_driver.F90:298-450 */
void fo_flux2dtracer(const fo_mesh *m, double dt, double *dttf_h, double *dttf_v,
                     const double *flux_h, const double *flux_v, int use_lo, const double *ttf,
                     const double *lo)
{
    const int L = m->nl - 1, NL = m->nl;
    if (use_lo)
        for (int n = 1; n <= m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz)
                dttf_v[IX2(L, nz, n)] = dttf_v[IX2(L, nz, n)] - ttf[IX2(L, nz, n)] * m->hnode[IX2(L, nz, n)] +
                                        lo[IX2(L, nz, n)] * m->hnode_new[IX2(L, nz, n)];
    for (int n = 1; n <= m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz)
            dttf_v[IX2(L, nz, n)] = dttf_v[IX2(L, nz, n)] +
                                    (flux_v[IX2(NL, nz, n)] - flux_v[IX2(NL, nz + 1, n)]) * dt / m->areasvol[IX2(NL, nz, n)];
    for (int edge = 1; edge <= m->edge2D; ++edge) {
        const int n1 = m->edges[IX2(2, 1, edge)], n2 = m->edges[IX2(2, 2, edge)];
        int nu12, nl12;
        edge_union_range(m, edge, &nu12, &nl12);
        for (int nz = nu12; nz <= nl12; ++nz) {
            dttf_h[IX2(L, nz, n1)] = dttf_h[IX2(L, nz, n1)] + flux_h[IX2(L, nz, edge)] * dt / m->areasvol[IX2(NL, nz, n1)];
            dttf_h[IX2(L, nz, n2)] = dttf_h[IX2(L, nz, n2)] - flux_h[IX2(L, nz, edge)] * dt / m->areasvol[IX2(NL, nz, n2)];
        }
    }
}

/* lib/oce_adv_tra_driver.F90:23-294 */
int fo_do_oce_adv_tra(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                      const double *w, const double *wi, const double *we, const double *ttf,
                      const double *ttfAB, fo_work *wk)
{
    const int L = m->nl - 1, NL = m->nl;
    const int fct = sc->lim == FO_LIM_FCT;
    if (fct) {
        fo_adv_tra_hor_upw1(m, vel, ttf, wk->adv_flux_hor, 1);
        double *lo = wk->fct_LO;
        memset(lo, 0, sizeof(double) * (size_t)L * (size_t)m->nod2D);
        for (int e = 1; e <= m->edge2D; ++e) {
            const int n1 = m->edges[IX2(2, 1, e)], n2 = m->edges[IX2(2, 2, e)];
            int nu12, nl12;
            edge_union_range(m, e, &nu12, &nl12);
            for (int nz = nu12; nz <= nl12; ++nz) {
                lo[IX2(L, nz, n1)] = lo[IX2(L, nz, n1)] + wk->adv_flux_hor[IX2(L, nz, e)];
                lo[IX2(L, nz, n2)] = lo[IX2(L, nz, n2)] - wk->adv_flux_hor[IX2(L, nz, e)];
            }
        }
        fo_adv_tra_ver_upw1(m, we, ttf, wk->adv_flux_ver, 1);
        for (int n = 1; n <= m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n - 1]; nz <= m->nlevels_nod2D[n - 1] - 1; ++nz)
                lo[IX2(L, nz, n)] = (ttf[IX2(L, nz, n)] * m->hnode[IX2(L, nz, n)] +
                                     (lo[IX2(L, nz, n)] + (wk->adv_flux_ver[IX2(NL, nz, n)] - wk->adv_flux_ver[IX2(NL, nz + 1, n)])) *
                                         dt / m->areasvol[IX2(NL, nz, n)]) /
                                    m->hnode_new[IX2(L, nz, n)];
        if (sc->use_wsplit) {
            /* :221-231 implicit part with wi, then the low-order vertical flux of the full w */
            if (fo_adv_tra_vert_impl(m, dt, wi, lo)) return -1;
            fo_adv_tra_ver_upw1(m, w, ttf, wk->adv_flux_ver, 1);
        }
    }
    const int zero = !fct;
    switch (sc->hor) {
    case FO_HOR_MUSCL:
        fo_adv_tra_hor_muscl(m, vel, ttfAB, sc->ph, wk->adv_flux_hor, wk->edge_up_dn_grad, zero);
        break;
    case FO_HOR_MFCT:
        fo_adv_tra_hor_mfct(m, vel, ttfAB, sc->ph, wk->adv_flux_hor, wk->edge_up_dn_grad, zero);
        break;
    case FO_HOR_UPW1:
        fo_adv_tra_hor_upw1(m, vel, ttfAB, wk->adv_flux_hor, zero);
        break;
    default:
        return -1;
    }
    const double *pw = fct ? w : we;
    switch (sc->ver) {
    case FO_VER_QR4C: fo_adv_tra_ver_qr4c(m, pw, ttfAB, sc->pv, wk->adv_flux_ver, zero); break;
    case FO_VER_CDIFF: fo_adv_tra_ver_cdiff(m, pw, ttfAB, wk->adv_flux_ver, zero); break;
    case FO_VER_PPM: fo_adv_tra_vert_ppm(m, dt, pw, ttfAB, wk->adv_flux_ver, zero); break;
    case FO_VER_UPW1: fo_adv_tra_ver_upw1(m, pw, ttfAB, wk->adv_flux_ver, zero); break;
    default: return -1;
    }
    if (fct) {
        fo_oce_tra_adv_fct(m, dt, ttf, wk->fct_LO, wk->adv_flux_hor, wk->adv_flux_ver, wk);
        fo_flux2dtracer(m, dt, wk->dttf_h, wk->dttf_v, wk->adv_flux_hor, wk->adv_flux_ver, 1, ttf,
                        wk->fct_LO);
    } else {
        fo_flux2dtracer(m, dt, wk->dttf_h, wk->dttf_v, wk->adv_flux_hor, wk->adv_flux_ver, 0, NULL,
                        NULL);
    }
    return 0;
}

/* Caller step (src/fesom_analytic.F90:404-437) with a given valuesAB. */
int fo_tracer_step_ab(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                      const double *w, const double *wi, const double *we, double *values,
                      const double *values_ab, fo_work *wk)
{
    const int L = m->nl - 1;
    memset(wk->dttf_h, 0, sizeof(double) * (size_t)L * (size_t)m->nod2D);
    memset(wk->dttf_v, 0, sizeof(double) * (size_t)L * (size_t)m->nod2D);
    if (sc->hor == FO_HOR_MUSCL || sc->hor == FO_HOR_MFCT) {
        fo_tracer_gradients(m, values_ab, wk->tr_xy);
        fo_fill_up_dn_grad(m, wk->tr_xy, wk->edge_up_dn_grad);
    }
    int rc = fo_do_oce_adv_tra(m, sc, dt, vel, w, wi, we, values, values_ab, wk);
    if (rc) return rc;
    for (int k = 1; k <= m->nod2D; ++k)
        for (int nz = m->ulevels_nod2D[k - 1]; nz <= m->nlevels_nod2D[k - 1] - 1; ++nz)
            values[IX2(L, nz, k)] = values[IX2(L, nz, k)] +
                                    (wk->dttf_h[IX2(L, nz, k)] + wk->dttf_v[IX2(L, nz, k)]) / m->hnode_new[IX2(L, nz, k)];
    return 0;
}

/* Caller step (src/fesom_analytic.F90:404-437 with valuesAB := values). */
int fo_tracer_step(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                   const double *w, const double *wi, const double *we, double *values,
                   fo_work *wk)
{
    const int L = m->nl - 1;
    memset(wk->dttf_h, 0, sizeof(double) * (size_t)L * (size_t)m->nod2D);
    memset(wk->dttf_v, 0, sizeof(double) * (size_t)L * (size_t)m->nod2D);
    if (sc->hor == FO_HOR_MUSCL || sc->hor == FO_HOR_MFCT) {
        fo_tracer_gradients(m, values, wk->tr_xy);
        fo_fill_up_dn_grad(m, wk->tr_xy, wk->edge_up_dn_grad);
    }
    int rc = fo_do_oce_adv_tra(m, sc, dt, vel, w, wi, we, values, values, wk);
    if (rc) return rc;
    for (int k = 1; k <= m->nod2D; ++k)
        for (int nz = m->ulevels_nod2D[k - 1]; nz <= m->nlevels_nod2D[k - 1] - 1; ++nz)
            values[IX2(L, nz, k)] = values[IX2(L, nz, k)] +
                                    (wk->dttf_h[IX2(L, nz, k)] + wk->dttf_v[IX2(L, nz, k)]) / m->hnode_new[IX2(L, nz, k)];
    return 0;
}

int fo_work_alloc(const fo_mesh *m, fo_work *wk)
{
    const size_t L = (size_t)(m->nl - 1), NL = (size_t)m->nl;
    const size_t N = (size_t)m->nod2D, T = (size_t)m->elem2D, E = (size_t)m->edge2D;
    memset(wk, 0, sizeof(*wk));
    wk->tr_xy = (double *)calloc(2 * L * T, sizeof(double));
    wk->edge_up_dn_grad = (double *)calloc(4 * L * (E > T ? E : T), sizeof(double));
    wk->fct_LO = (double *)calloc(L * N, sizeof(double));
    wk->adv_flux_hor = (double *)calloc(L * E, sizeof(double));
    wk->adv_flux_ver = (double *)calloc(NL * N, sizeof(double));
    wk->fct_ttf_max = (double *)calloc(L * N, sizeof(double));
    wk->fct_ttf_min = (double *)calloc(L * N, sizeof(double));
    wk->fct_plus = (double *)calloc(L * N, sizeof(double));
    wk->fct_minus = (double *)calloc(L * N, sizeof(double));
    wk->tvert_max = (double *)calloc(L * N, sizeof(double));
    wk->tvert_min = (double *)calloc(L * N, sizeof(double));
    wk->dttf_h = (double *)calloc(L * N, sizeof(double));
    wk->dttf_v = (double *)calloc(L * N, sizeof(double));
    if (!wk->tr_xy || !wk->edge_up_dn_grad || !wk->fct_LO || !wk->adv_flux_hor || !wk->adv_flux_ver ||
        !wk->fct_ttf_max || !wk->fct_ttf_min || !wk->fct_plus || !wk->fct_minus || !wk->tvert_max ||
        !wk->tvert_min || !wk->dttf_h || !wk->dttf_v) {
        fo_work_free(wk);
        return -1;
    }
    return 0;
}

void fo_work_free(fo_work *wk)
{
    free(wk->tr_xy);
    free(wk->edge_up_dn_grad);
    free(wk->fct_LO);
    free(wk->adv_flux_hor);
    free(wk->adv_flux_ver);
    free(wk->fct_ttf_max);
    free(wk->fct_ttf_min);
    free(wk->fct_plus);
    free(wk->fct_minus);
    free(wk->tvert_max);
    free(wk->tvert_min);
    free(wk->dttf_h);
    free(wk->dttf_v);
    memset(wk, 0, sizeof(*wk));
}
