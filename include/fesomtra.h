/*
 * fesomtra: MI355X (gfx950) tracer advection for FESOM2 meshes -- C ABI.
 *
 * Drop-in for the cross-language interface of sbrdar/fesom2_tracer_dwarf
 * (lib/tracer_c_interface.F90): same symbol names, argument meaning, int status codes
 * (0 = OK, -1 = state or argument error, message on stderr), Fortran column-major arrays with
 * 1-based index values, copy-in / copy-out semantics and one process-wide instance.  Unlike the
 * reference, the library owns device (HBM) copies of everything; the horizontal, vertical and
 * limiter stages of do_oce_adv_tra (lib/oce_adv_tra_driver.F90:23-294) run as HIP kernels.
 *
 * Conventions (as the reference): a(d1, d2) means a Fortran array, i.e. a[(j-1)*d1 + (i-1)];
 * node / element / edge numbers are 1-based; "nl" is the number of levels (interfaces), so
 * tracers have nl-1 layers.
 *
 * Entry points marked [new] have no counterpart in the reference's C interface: the reference
 * can only advect after a restart read, so the connectivity, geometry and scheme setters that a
 * mesh needs are added here (they fill what lib/MOD_MESH.F90 T_MESH and lib/MOD_TRACER.F90
 * hold).  Not thread-safe; call from one thread.
 */
#ifndef FESOMTRA_H
#define FESOMTRA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference entry points (lib/tracer_c_interface.F90) ---------------------------------- */

/* :46  Records the communicator; one GPU, one partition: no MPI is initialised. */
int tracer_init_mpi(int mpi_comm);
/* :89 */
int tracer_finalize_mpi(void);
/* :113 Reads the derived-type restart files <restart_dir>/np1/t_mesh.0, t_partit.0, t_tracer.0
 *      and t_dynamics.0 (lib/io_restart_derivedtype.F90, as the reference's FP64 build writes
 *      them: one unformatted record per type) and fills mesh, velocities (w, w_i, w_e, use_wsplit),
 *      tracers and their schemes through the setters; one partition (eDim = 0), no cavities.
 *      -1 "could not find restart directory" when <restart_dir>/np1 is missing. */
int tracer_init(const char *restart_dir);
/* :163 One time step of every tracer: zero tendencies, do_oce_adv_tra, values += (dttf_h +
 *      dttf_v)/hnode_new, with valuesAB := values.  (The reference advects tracer 1 only and
 *      never zeroes del_ttf; both are corrected here.)  Returns after the device finished. */
int tracer_advect_step(double dt);
/* :228 min / max / plain sum over all (nl-1) x nod2D entries of tracer 1. */
int tracer_get_stats(double *tmin, double *tmax, double *tsum);
/* :255 */
int tracer_get_size(int *nz, int *nn);
/* :280 Copies tracer 1 out, values(nz=nl-1, nn=nod2D). */
int tracer_get_values(double *values, int nz, int nn);
/* :308 Writes <output_dir>/np1/t_*.0 and <output_dir>/meta.time in the reference's restart
 *      format with the current tracers (and their valuesAB), schemes and velocities; a state read
 *      by tracer_init keeps every other item of its files, a state built from arrays writes the
 *      items the advection reads (the rest as not allocated).  Ends the initialised state, as the
 *      reference does. */
int tracer_finalize(const char *output_dir);
/* :356 tracer_init(restart_dir), nsteps x (tracer_advect_step, tracer_get_stats, a line
 *      "Step n: min/max/sum = ..."), tracer_finalize(restart_dir). */
int tracer_run_workflow(const char *restart_dir, int nsteps, double dt);
/* :391 Releases all device and host state. */
int tracer_cleanup(void);
/* :415 */
int tracer_set_mesh_dims(int nl, int nod2D, int edge2D, int elem2D);
/* :441 nlevels_nod2D, ulevels_nod2D (nod2D). */
int tracer_set_levels(const int *nlevels, const int *ulevels, int nod2D);
/* :477 hnode(nl-1, nod2D); sets hnode and hnode_new (the reference sets hnode_new only). */
int tracer_set_thickness(const double *hnode, int nl, int nod2D);
/* :510 eDim_nod2D must be 0 (single partition; the periodic wrap lives in the connectivity). */
int tracer_set_partition(int myDim_nod2D, int eDim_nod2D);
/* :539 Schemes default to UPW1/UPW1/NONE, ph = pv = 1 (as the reference). */
int tracer_allocate_tracers(int num_tracers, int nl, int nod2D, int AB_order);
/* :613 values(nl-1, nod2D), tracer_id 1-based. */
int tracer_set_values(int tracer_id, const double *values, int nl, int nod2D);
/* :652 uv, w, w_i, w_e := 0.  edge2D only sized the reference's own edge array and is not checked. */
int tracer_set_velocity_zero(int nl, int nod2D, int edge2D);
/* :694 The reference stores uv as (nl-1, edge2D, 2) although the kernels read element velocities
 *      vel(2, nl-1, elem2D); this entry point accepts only an all-zero uv (sets w) and otherwise
 *      returns -1 pointing to tracer_set_velocity_elem. */
int tracer_set_velocity(const double *uv, const double *w, int nl, int nod2D, int edge2D);
/* :749 Validates the mesh, reorders it along a space-filling curve and uploads it to HBM. */
int tracer_init_complete(void);
/* :769 As tracer_init without the tracers: mesh and velocities from <restart_dir>/np1/t_mesh.0,
 *      t_partit.0 and t_dynamics.0; tracers follow with tracer_allocate_tracers. */
int tracer_load_mesh_partition(const char *restart_dir);
/* :830 Reads a FESOM2 mesh directory (nod2d.out, elem2d.out, aux3d.out, nlvls.out, elvls.out,
 *      edgenum.out, edges.out, edge_tri.out; one partition, no cavities), rotates the nodes to
 *      the model grid (default Euler angles 50, 15, -90 degrees), orients the triangles and
 *      derives the spherical geometry and the MUSCL support arrays as the reference's mesh set-up
 *      does (longitude differences wrapped into [-pi, pi]); replaces tracer_set_mesh_dims and the
 *      mesh setters.  Tracers, schemes and element velocities follow as with an array mesh.  As
 *      in the reference, the loaded mesh counts as initialised: tracer_get_size,
 *      tracer_init_ale_arrays, tracer_init_for_advection and stepping work before
 *      tracer_init_complete (velocities default to zero). */
int tracer_load_mesh_from_files(const char *mesh_path);
/* :875 hnode = hnode_new = helem = zbar(k) - zbar(k+1) from tracer_set_zbar. */
int tracer_init_ale_arrays(void);
/* :931 Allocates the advection work arrays (done lazily otherwise). */
int tracer_init_for_advection(void);

/* ---- [new] mesh, dynamics and scheme setters ------------------------------------------------ */

/* coord_nod2D(2, nod2D) in radians (lib/MOD_MESH.F90 coord_nod2D); used for the node order. */
int tracer_set_coordinates(const double *coord_nod2D, int nod2D);
/* elem2D_nodes(3, elem2D) clockwise, edges(2, edge2D), edge_tri(2, edge2D) with edge_tri(1) the
 * element left of edges(1)->edges(2) and 0 for none, nod_in_elem2D(max_adj, nod2D),
 * nod_in_elem2D_num(nod2D). */
int tracer_set_connectivity(const int *elem2D_nodes, const int *edges, const int *edge_tri,
                            const int *nod_in_elem2D, const int *nod_in_elem2D_num, int max_adj,
                            int nod2D, int elem2D, int edge2D);
/* nlevels, ulevels of elements (elem2D). */
int tracer_set_elem_levels(const int *nlevels, const int *ulevels, int elem2D);
/* elem_area(elem2D), area(nl, nod2D), areasvol(nl, nod2D), edge_dxdy(2, edge2D) [radians],
 * edge_cross_dxdy(4, edge2D) [m], elem_cos(elem2D), gradient_sca(6, elem2D)
 * (lib/oce_mesh.F90 mesh_areas / mesh_auxiliary_arrays). */
int tracer_set_geometry(const double *elem_area, const double *area, const double *areasvol,
                        const double *edge_dxdy, const double *edge_cross_dxdy,
                        const double *elem_cos, const double *gradient_sca, int nl, int nod2D,
                        int elem2D, int edge2D);
/* hnode, hnode_new, Z_3d_n (nl-1, nod2D), helem (nl-1, elem2D), zbar_3d_n (nl, nod2D). */
int tracer_set_vertical_grid(const double *hnode, const double *hnode_new, const double *helem,
                             const double *Z_3d_n, const double *zbar_3d_n, int nl, int nod2D,
                             int elem2D);
/* zbar(nl): reference interface depths, for tracer_init_ale_arrays. */
int tracer_set_zbar(const double *zbar, int nl);
/* MUSCL support (lib/oce_muscl_adv.F90): edge_up_dn_tri(2, edge2D), nboundary_lay(nod2D).
 * Without it the library derives nboundary_lay itself and MUSCL/MFCT are unavailable. */
int tracer_set_muscl_support(const int *edge_up_dn_tri, const int *nboundary_lay, int nod2D,
                             int edge2D);
/* Element velocities vel(2, nl-1, elem2D) and w, w_i, w_e (nl, nod2D), the arrays
 * do_oce_adv_tra reads (lib/oce_adv_tra_driver.F90:42-45).  w_i / w_e may be NULL (= w). */
int tracer_set_velocity_elem(const double *uv, const double *w, const double *w_i,
                             const double *w_e, int nl, int nod2D, int elem2D);
/* [new] Time-step factor per node, factor(nod2D) > 0 (NULL: 1 everywhere, the default): the
 * cells of node n advance with dt * factor(n), where dt is the argument of the stepping calls.
 * For a mesh that is a disjoint union of independent meshes (the library advances them in one
 * launch sequence) each member keeps its own time step: pass its dt as the factor of its nodes
 * and step with dt = 1.  A factor that differs between the two ends of an edge is not
 * conservative; the library does not check it. */
int tracer_set_node_dt(const double *factor, int nod2D);
/* [new] Ensemble members: nmembers (1 to 64, default 1) velocity fields on the one mesh, each
 * tracer advected by the velocity of its member.  The reference runs one velocity per model
 * instance (do_oce_adv_tra takes vel, w, w_i, w_e per call, lib/oce_adv_tra_driver.F90:23-45);
 * here the members of a tracer block share every mesh read and each lane reads its member's
 * transport and vertical velocity.  Member 1 is the velocity of tracer_set_velocity_elem; new
 * members start as a copy of it.  Changing the count rebuilds the device copy (tracer values are
 * kept); restart files (tracer_finalize) are refused for more than one member. */
int tracer_set_ensemble(int nmembers);
int tracer_get_ensemble(int *nmembers);
/* [new] Velocity of member `member` (1-based, layouts of tracer_set_velocity_elem; member 1 is
 * tracer_set_velocity_elem itself). */
int tracer_set_velocity_member(int member, const double *uv, const double *w, const double *w_i,
                               const double *w_e, int nl, int nod2D, int elem2D);
/* [new] Member (1-based) whose velocity advects tracer `tracer_id` (default 1). */
int tracer_set_tracer_member(int tracer_id, int member);
/* Scheme of one tracer with the reference's names: hor "UPW1" | "MUSCL" | "MFCT",
 * ver "UPW1" | "QR4C" | "CDIFF" | "PPM", lim "NONE" | "FCT"; ph, pv as tra_adv_ph / tra_adv_pv
 * (lib/MOD_TRACER.F90:18-22).  Unknown names return -1 (the reference aborts). */
int tracer_set_scheme(int tracer_id, const char *hor, const char *ver, const char *lim,
                      double ph, double pv);

/* ---- [new] stepping, drop-in and inspection ------------------------------------------------- */

/* nsteps calls of tracer_advect_step without host synchronisation in between. */
int tracer_advect_steps(int nsteps, double dt);
/* The do_oce_adv_tra contract for one tracer (lib/oce_adv_tra_driver.F90:23-294): accumulates
 * into the device tendencies del_ttf_advhoriz / del_ttf_advvert (zeroed by
 * tracer_zero_tendencies, as the caller zeroes them) from values and valuesAB of tracer_id.
 * The edge transport of the velocity is computed once per velocity (tracer_set_velocity_elem
 * starts a new one) and shared by the tracers; the kernels are queued on the library's stream
 * and the call returns without waiting for them (every call that reads results waits). */
int tracer_do_oce_adv_tra(double dt, int tracer_id);
int tracer_zero_tendencies(void);
/* Copies del_ttf_advhoriz / del_ttf_advvert (nz=nl-1, nn=nod2D) out. */
int tracer_get_tendencies(double *dttf_h, double *dttf_v, int nz, int nn);
/* valuesAB of one tracer (nl-1, nod2D), for tracer_do_oce_adv_tra; tracer_set_values and
 * tracer_advect_step reset it to values. */
int tracer_set_values_ab(int tracer_id, const double *values, int nl, int nod2D);
/* Device-pointer variants of tracer_set_values_ab and tracer_get_tendencies for callers whose
 * arrays live in HBM: valuesAB(nl-1, nod2D) and dttf_h / dttf_v (nz=nl-1, nn=nod2D), float64 in
 * reference order (either tendency pointer may be NULL). */
int tracer_set_values_ab_device(int tracer_id, const void *values, int nl, int nod2D);
int tracer_get_tendencies_device(void *dttf_h, void *dttf_v, int nz, int nn);
/* [new, no counterpart in the reference] tracer_do_oce_adv_tra for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): values and valuesAB in,
 * dttf_h / dttf_v accumulated (+=, as do_oce_adv_tra does), all device float64 (nl-1, nod2D) in
 * reference order; every work array of the pass lives in a fixed 256 MiB device array of the
 * library (one per GPU, held until the process ends).  Allocates nothing and waits for the
 * kernels before returning.  Needs FP64 (tracer_set_precision(0)) and the mesh on the device
 * (tracer_init_for_advection or any step before); returns -1 with a message for a mesh whose
 * work arrays exceed the fixed array: 7 x (nl-1) x nod2D + nl x nod2D + (nl-1) x edge2D doubles
 * (2 x (nl-1) x nod2D more with use_wsplit), about 2.5 M node-levels of a triangular mesh.  The
 * array keeps the last call's contents and nothing in it needs to start at zero; calls that
 * overlap on one GPU race on it. */
int tracer_do_oce_adv_tra_noalloc(double dt, int tracer_id, const void *values, const void *values_ab,
                                  void *dttf_h, void *dttf_v);
int tracer_get_values_by_id(int tracer_id, double *values, int nz, int nn);
/* Over the wet layers of tracer_id: min, max, plain sum and sum of T*hnode*areasvol (FP64). */
int tracer_get_stats_by_id(int tracer_id, double *tmin, double *tmax, double *tsum,
                           double *vsum);
/* Work arrays of the last tracer_do_oce_adv_tra call for inspection: name is "fct_LO"
 * (nl-1, nod2D), "adv_flux_hor" (nl-1, edge2D) or "adv_flux_ver" (nl, nod2D), the latter two
 * after limiting as the reference leaves them.  count = number of doubles.  Returns -1 when a
 * tracer_advect_step(s) call came after it (the steps reuse the arrays for every tracer).  In the
 * single-precision modes "fct_LO" holds the increment fct_LO - values, and "adv_flux_hor" of a
 * tracer without the limiter is not available (-1: the kernels store the high-order minus the
 * upwind flux). */
int tracer_get_work(const char *name, double *out, size_t count);
/* Whether tracer_do_oce_adv_tra keeps the limited fluxes "adv_flux_hor" / "adv_flux_ver" for
 * tracer_get_work (keep = 1, the default, as the reference leaves them in its work arrays), or
 * only accumulates the tendencies (keep = 0: 0.9 GB fewer stores per call on a 1000x1000x48 mesh;
 * tracer_get_work then returns -1 for those two arrays; "fct_LO" is always kept). */
int tracer_set_keep_work(int keep);
/* Fused limiter (enable = 1; default 0): the limiter bounds and ratios, the limiting and the
 * update of FP64 tracer blocks of width <= 4 run as one kernel on tiles of 128 / width nodes whose
 * first ring has threads of its own, which form that ring's ratios themselves instead of storing
 * and re-reading them; the values then alternate between two buffers per step (the stored values
 * are always the primary buffer's after a call).  Results are bit-identical either way.  It
 * applies when every pass has the FCT limiter, there are no cavities, per-node time steps,
 * ensembles or AB2 time levels and every tile's two rings fit its threads and staging slots;
 * tracer_get_fused_limiter reports whether the next steps use it (1) or the two kernels (0).  On
 * MI355X it is slower than the two kernels (1000x1000x48 mesh: 1.36 against 1.26 ms per
 * tracer-step at width 1, 2.62 against 2.17 ms for a block of two: the first ring adds about half
 * again to the ratio work, and the fused state leaves three waves per SIMD), so it is off by
 * default. */
/* Concurrent low-order kernels (enable = 1, the default): with the FCT limiter the low-order solution
 * (and the implicit split) runs on a second stream of the library beside the gradient and
 * edge-flux kernels, which it does not depend on, joined before the limiter bounds; 0 runs every
 * kernel on one stream.  Results are identical either way; graphs capture both streams. */
int tracer_set_concurrent(int enable);
int tracer_set_fused_limiter(int enable);
int tracer_get_fused_limiter(int *active);
/* Device-pointer variants of tracer_set_values / tracer_get_values for callers whose tracers
 * already live in HBM (e.g. PyTorch tensors): values(nl-1, nod2D) float64 in reference order. */
int tracer_set_values_device(int tracer_id, const void *values, int nl, int nod2D);
int tracer_get_values_device(int tracer_id, void *values, int nz, int nn);
/* Wall time of the last tracer_advect_steps call measured with HIP events, in ms. */
int tracer_last_elapsed_ms(double *ms);
/* Per-stage HIP-event timing of the advect calls that follow (enable = 1) on the library's
 * stream: stage 0 edge fluxes, 1 low-order solution / vertical fluxes, 2 limiter bounds, 3
 * limiting + update.  tracer_get_kernel_times returns the mean ms per tracer-step of each stage
 * and the number of tracer-steps averaged. */
int tracer_set_kernel_timing(int enable);
int tracer_get_kernel_times(double *stage_ms, long *count);
/* With the kernel timing on: mean ms per step of the edge-transport kernel (k_qflux, once per
 * step) and of the whole step (from before k_qflux to after the last pass), and the number of
 * steps averaged. */
int tracer_get_step_times(double *qflux_ms, double *step_ms, long *count);
/* Arithmetic / storage precision (applies to all tracers; values are carried over):
 * 0 FP64 everywhere (default; bit-identical to the reference dp build),
 * 1 FP32 flux and reconstruction arithmetic, FP32 mesh and tracer storage,
 * 2 as 1 with FP16 tracer storage, 3 as 1 with BF16 tracer storage.
 * In modes 1-3 the nodal flux-divergence sums, the low-order solution, the limiter ratios, the
 * tracer update and the global sums stay in FP64. */
int tracer_set_precision(int mode);
/* Rounding of the tracer update into single / half / bfloat16 storage (modes 1-3): 1 =
 * stochastic (default; rounds up with probability equal to the remainder, from a counter-based
 * hash of tracer number, cell and step, so results are reproducible and independent of the block
 * width), 0 = round to nearest.  Without it, increments smaller than half a unit in the last place
 * are lost at every step and slowly varying fields stagnate. */
int tracer_set_rounding(int stochastic);
/* Compensated storage (modes 1-3; enable = 1): next to every stored value the library keeps its
 * rounding remainder (value - stored, in FP32), and next to the stored low-order increment its own
 * remainder; the update adds the step's increment to value + remainder in double and keeps the new
 * remainder, so no storage rounding accumulates: the tracer content sum(T * hnode * areasvol)
 * changes only by the (conservative) fluxes, to FP64 round-off, and values, statistics and
 * read-backs are stored value + remainder.  The fluxes are computed from the stored values.  Costs
 * one more FP32 read and write per node-level and step for the values and one for the increment.
 * Default 0; values are carried over when it changes. */
int tracer_set_compensated(int enable);
/* restore = 0: keeps a device copy of every tracer (in HBM, storage precision) and the step
 * counter of the stochastic rounding; restore = 1: puts both back (valuesAB := values).  Lets a
 * caller time or tune steps without changing later results. */
int tracer_checkpoint(int restore);
/* Tracer block width kb (a power of two up to 64; 0 = automatic, the default: all tracers, up to
 * 64, in one block).  Tracer id i is stored as lane (i-1) % kb of block (i-1) / kb, interleaved
 * innermost in HBM, and one pass of the four kernels advances every lane of a block that shares
 * a scheme, so the mesh and velocities are read once for all of them.  Stored values are carried
 * over when the width changes; results do not depend on it. */
int tracer_set_tracers_per_pass(int kb);
int tracer_get_tracers_per_pass(int *kb);
/* enable = 1: tracer_advect_steps(nsteps, dt) captures its whole launch sequence (nsteps x the
 * passes of every block) into one HIP graph at the first call and replays it for later calls
 * with the same nsteps and dt until the scheme, precision, width or launch configuration
 * changes.  Per-stage kernel timing, when enabled, uses the uncaptured path. */
int tracer_set_graph(int enable);
/* enable = 1: FP64 tracers with the FCT limiter in blocks of width 1 are advanced by one fused
 * kernel per step (tiles of nodes with a three-ring halo, exchanging values level by level in
 * LDS), which stores no intermediate arrays.  Results are bit-identical to the default path.
 * Returns -1 if the mesh cannot be tiled (a node with more than 8 edges); other tracers keep the
 * multi-kernel path. */
int tracer_set_fused(int enable);
/* dynamics%use_wsplit of do_oce_adv_tra (lib/oce_adv_tra_driver.F90:221-231): enable = 1 splits
 * the vertical velocity for tracers with the FCT limiter: the low-order solution takes the
 * explicit upwind flux of w_e, then the implicit first-order upwind advection with w_i
 * (adv_tra_vert_impl, lib/oce_adv_tra_ver.F90:90-244, a tridiagonal solve per column), and the
 * low-order part of the antidiffusive vertical flux is the upwind flux of the full w.  w, w_i and
 * w_e come from tracer_set_velocity_elem.  Default 0, as the reference. */
int tracer_set_wsplit(int enable);
int tracer_get_wsplit(int *enabled);
/* valuesAB, the time level the high-order fluxes and gradients read, in tracer_advect_step(s):
 * mode 0 (default): valuesAB := values at every step.
 * mode 1: frozen, as the reference's drivers leave it: set to values by tracer_set_values (and
 *         when the mode is selected) and never changed by a step (src/fesom_analytic.F90:335,
 *         lib/tracer_c_interface.F90:613-648, lib/tracer_c_interface.F90:163-225).
 * mode 2: FESOM2's second-order Adams-Bashforth interpolation before every step,
 *         valuesAB = -(0.5+epsilon)*valuesold + (1.5+epsilon)*values, then valuesold := values
 *         (epsilon = 0.1 in lib/oce_modules.F90:75, valuesold lib/MOD_TRACER.F90:13); selecting
 *         the mode and tracer_set_values set valuesold := values.
 * epsilon in [0, 1] (used by mode 2).  valuesAB / valuesold are held at the storage precision and
 * kept across changes of velocity, precision or block width; tracer_checkpoint keeps them too. */
int tracer_set_ab_mode(int mode, double epsilon);
int tracer_get_ab_mode(int *mode, double *epsilon);
/* enable = 1 (default): for tracer blocks of width 1, 2 or 4 the kernels run on tiles that stage
 * neighbour values in LDS, level by level, instead of gathering them once per edge: the edge-flux
 * kernel on tiles of 256 / width consecutive edges (with the nodes they read), and with the FCT
 * limiter the low-order, bounds and limiting kernels on tiles of 256 / width consecutive nodes
 * (a compact patch of the renumbered mesh) with their one-ring halo.  Results are bit-identical
 * either way; meshes whose tiles need more than the staging space use the untiled kernels.
 * tracer_get_halo_tiles returns the number of node tiles in use (0: untiled). */
int tracer_set_halo_tiles(int enable);
int tracer_get_halo_tiles(int *ntiles);
/* Levels per thread of the four kernels (edge fluxes, low-order / vertical, bounds, final): each
 * column is split into chunks of that many layers, one thread per chunk and tracer lane.  The
 * default is 16 for each; results do not depend on the choice. */
int tracer_set_launch_config(int lc_edge, int lc_lo, int lc_bounds, int lc_final);
int tracer_get_launch_config(int *cfg4);
/* Last error message (empty if none). */
const char *tracer_last_error(void);
/* Library version string. */
const char *tracer_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FESOMTRA_H */
