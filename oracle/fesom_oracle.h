/* CPU restatement of the FESOM2 tracer advection step (sbrdar/fesom2_tracer_dwarf, the
 * do_oce_adv_tra path).  TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  Never linked into or called by
 * the shipped GPU library.
 *
 * Arrays keep the reference's Fortran layout and 1-based index values: an array declared
 * a(k, n) in the reference is a flat double[] with a[(n-1)*ld + (k-1)].  Loops and floating-point
 * operation order follow the reference routines line by line (citations at each function), so a
 * build with -ffp-contract=off reproduces the serial Fortran dp build bit for bit.
 */
#ifndef FESOM_ORACLE_H
#define FESOM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_mesh {
    int nl, nod2D, elem2D, edge2D, max_adj;
    const int *elem2D_nodes;       /* (3, elem2D) */
    const int *edges;              /* (2, edge2D) */
    const int *edge_tri;           /* (2, edge2D), 0 = none */
    const int *nod_in_elem2D;      /* (max_adj, nod2D) */
    const int *nod_in_elem2D_num;  /* (nod2D) */
    const int *nlevels, *ulevels;  /* (elem2D) */
    const int *nlevels_nod2D, *ulevels_nod2D;          /* (nod2D) */
    const int *nlevels_nod2D_min, *ulevels_nod2D_max;  /* (nod2D) */
    const int *edge_up_dn_tri;     /* (2, edge2D) */
    const int *nboundary_lay;      /* (nod2D) */
    const double *elem_area;       /* (elem2D) */
    const double *area, *areasvol; /* (nl, nod2D) */
    const double *edge_dxdy;       /* (2, edge2D) radians */
    const double *edge_cross_dxdy; /* (4, edge2D) metres */
    const double *elem_cos;        /* (elem2D) */
    const double *gradient_sca;    /* (6, elem2D) */
    const double *hnode, *hnode_new; /* (nl-1, nod2D) */
    const double *helem;           /* (nl-1, elem2D) */
    const double *Z_3d_n;          /* (nl-1, nod2D) */
    const double *zbar_3d_n;       /* (nl, nod2D) */
} fo_mesh;

/* Scheme codes, same names as the reference's strings. */
enum { FO_HOR_UPW1 = 0, FO_HOR_MUSCL = 1, FO_HOR_MFCT = 2 };
enum { FO_VER_UPW1 = 0, FO_VER_QR4C = 1, FO_VER_CDIFF = 2, FO_VER_PPM = 3 };
enum { FO_LIM_NONE = 0, FO_LIM_FCT = 1 };

typedef struct fo_scheme {
    int hor, ver, lim;
    double ph, pv;
    int use_wsplit; /* dynamics%use_wsplit: implicit low-order vertical part with wi */
} fo_scheme;

/* Work arrays of one tracer step (the reference's T_TRACER_WORK plus tr_xy). */
typedef struct fo_work {
    double *tr_xy;            /* (2, nl-1, elem2D) */
    double *edge_up_dn_grad;  /* (4, nl-1, edge2D) */
    double *fct_LO;           /* (nl-1, nod2D) */
    double *adv_flux_hor;     /* (nl-1, edge2D) */
    double *adv_flux_ver;     /* (nl, nod2D) */
    double *fct_ttf_max, *fct_ttf_min, *fct_plus, *fct_minus; /* (nl-1, nod2D) */
    double *tvert_max, *tvert_min;                             /* (nl-1, nod2D) */
    double *dttf_h, *dttf_v;  /* (nl-1, nod2D) */
} fo_work;

/* Individual routines (arguments as in the reference; init_zero as o_init_zero). */
void fo_tracer_gradients(const fo_mesh *m, const double *ttf, double *tr_xy);
void fo_fill_up_dn_grad(const fo_mesh *m, const double *tr_xy, double *edge_up_dn_grad);
void fo_adv_tra_hor_upw1(const fo_mesh *m, const double *vel, const double *ttf, double *flux,
                         int init_zero);
void fo_adv_tra_hor_muscl(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                          double *flux, const double *edge_up_dn_grad, int init_zero);
void fo_adv_tra_hor_mfct(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                         double *flux, const double *edge_up_dn_grad, int init_zero);
void fo_adv_tra_ver_upw1(const fo_mesh *m, const double *w, const double *ttf, double *flux,
                         int init_zero);
void fo_adv_tra_ver_qr4c(const fo_mesh *m, const double *w, const double *ttf, double num_ord,
                         double *flux, int init_zero);
void fo_adv_tra_vert_ppm(const fo_mesh *m, double dt, const double *w, const double *ttf,
                         double *flux, int init_zero);
void fo_adv_tra_ver_cdiff(const fo_mesh *m, const double *w, const double *ttf, double *flux,
                          int init_zero);
void fo_oce_tra_adv_fct(const fo_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, fo_work *wk);
int fo_adv_tra_vert_impl(const fo_mesh *m, double dt, const double *w, double *ttf);
void fo_flux2dtracer(const fo_mesh *m, double dt, double *dttf_h, double *dttf_v,
                     const double *flux_h, const double *flux_v, int use_lo, const double *ttf,
                     const double *lo);

/* do_oce_adv_tra: accumulates into wk->dttf_h / wk->dttf_v (caller zeroes them).
 * edge_up_dn_grad must be filled before the call (it is clobbered, as in the reference). */
int fo_do_oce_adv_tra(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                      const double *w, const double *wi, const double *we, const double *ttf,
                      const double *ttfAB, fo_work *wk);

/* One caller step for one tracer: zero tendencies, valuesAB := values, gradients,
 * fill_up_dn_grad, do_oce_adv_tra, values += (dttf_h + dttf_v)/hnode_new. */
int fo_tracer_step(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                   const double *w, const double *wi, const double *we, double *values,
                   fo_work *wk);

/* The same with a separate valuesAB (gradients and high-order fluxes read values_ab). */
int fo_tracer_step_ab(const fo_mesh *m, const fo_scheme *sc, double dt, const double *vel,
                      const double *w, const double *wi, const double *we, double *values,
                      const double *values_ab, fo_work *wk);

/* Allocation helpers for fo_work (returns 0 on success). */
int fo_work_alloc(const fo_mesh *m, fo_work *wk);
void fo_work_free(fo_work *wk);

#ifdef __cplusplus
}
#endif
#endif
