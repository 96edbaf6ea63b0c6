/*
 * wmpic.h -- C ABI of libwmpic, the MI355X-native implementation of the
 * WumingPIC2D per-step hot path (push, deposit, sort, boundaries, implicit
 * field solve).
 *
 * Conventions
 *  - Every function returns an int status: WM_OK (0) or one of WM_ERR_*.
 *    wm_last_error() returns the message; for the reference's fatal stops
 *    the message is the reference's own text.
 *  - Arrays passed by host pointer use the reference's column-major layout
 *    and index bounds (first index fastest), exactly as the Fortran callers
 *    hold them:
 *      up, gp  (ndim, np, nys:nye, nsp)        real(8)
 *      uf, df  (6, nxgs-2:nxge+2, nys-2:nye+2)  real(8)
 *      uj      (3, nxgs-2:nxge+2, nys-2:nye+2)  real(8)
 *      cumcnt  (nxgs:nxge+1, nys:nye, nsp)      integer(4)
 *      np2     (nys:nye, nsp)                   integer(4)
 *      phi     (nxs-1:nxe+1, nys-1:nye+1)       real(8)
 *  - One context per device and grid (one MPI rank of the reference).  All
 *    calls on a context are made from one host thread; each call is
 *    synchronous with respect to the host-visible results it returns.
 *  - The host-array entry points (wm_particle_solv ... wm_bcp_phi) keep the
 *    reference's call semantics: inputs are read from host memory, outputs are
 *    written to host memory, persistent state (the CG warm start of
 *    field__fdtd_i) stays on the device.  The device-resident entry points
 *    (wm_state_*, wm_step) keep particles and fields in HBM across steps.
 *    The host-array entry points that stage particles or fields
 *    (particle_solv, field_fdtd_i, bcp_particle_x/y, sort_bucket,
 *    mom_calc_accl/nvt) use the buffers a device-resident state lives in: on
 *    a context that holds one (after wm_state_upload or wm_state_init_weibel)
 *    they return WM_ERR_ARG and change nothing.
 *
 * Citations name the reference routine each entry point replaces
 * (files of WumingCode/WumingPIC2D).
 */
#ifndef WMPIC_H
#define WMPIC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WM_OK 0
#define WM_ERR_NOT_INIT 1   /* "Initialize first by calling ..." particle.f90:63-66,
                               field.f90:100-103, sort.f90:47-50,
                               boundary_periodic.f90:69-72 */
#define WM_ERR_CG_ITEMAX 2  /* "stop at cgm after ite_max" field.f90:427-430 */
#define WM_ERR_NP_OVER 3    /* "memory over (np2 > np)" boundary_periodic.f90:231-234 */
#define WM_ERR_RANGE 4      /* a particle outside the cells the reference indexes */
#define WM_ERR_ARG 5        /* inconsistent arguments */
#define WM_ERR_HIP 6        /* HIP runtime error */
#define WM_ERR_NOMEM 7      /* device allocation failed */
#define WM_ERR_NODEV 8      /* no usable GPU */

/* Particle storage forms of the device-resident state (wm_set_form). */
#define WM_FORM_FP64 0      /* absolute x, y and u in FP64, the reference's arithmetic */
#define WM_FORM_FP32 1      /* int32 cell index + FP32 in-cell offsets + FP32 u */

/* Implicit-solve methods (wm_set_solver). */
#define WM_SOLVER_CG 0      /* the reference's plain CG (field.f90:319-461) */
#define WM_SOLVER_FCG 1     /* flexible CG with the learned preconditioner */

/* Boundary families for wm_field_fdtd_i's set_boundary_* callbacks. */
#define WM_BC_PERIODIC 0    /* boundary_periodic__dfield/__curre/__phi */
#define WM_BC_SHOCK 1       /* boundary_shock__*: conducting wall at nxs, open edge at nxe */
#define WM_BC_RECONNECTION 2 /* boundary_reconnection__*: conducting walls at nxs, nxe */

typedef struct wm_ctx wm_ctx;

/* ---- context (no reference counterpart: holds what the reference keeps in
 *      module SAVE variables, on the device) ---- */
int wm_create(wm_ctx **ctx, int device);
int wm_destroy(wm_ctx *ctx);
const char *wm_last_error(const wm_ctx *ctx);
/* Device capacity per (species, row) and the migration buffer.  The
 * reference sizes rows by np (proj/weibel/app.f90:248); the device may use
 * less.  mover_cap: FP64 form, migrating particles per row and direction;
 * FP32 form, mover records per row (all species) per push.  0 selects
 * defaults (np; np/8 or a quarter of a row's slots). */
int wm_set_capacity(wm_ctx *ctx, int cap_per_row, int mover_cap);
/* CG relative tolerance; the reference fixes 1e-6 (field.f90:339). */
int wm_set_cg_tolerance(wm_ctx *ctx, double err);
/* Particle form of the device-resident run (WM_FORM_FP64 default).  In the
 * FP32 form the push and the per-tile deposit run in FP32 while the current
 * accumulated across tiles, the field solve (CG dots and norms) and the
 * energy sums stay FP64; the host-array entry points need the FP64 form.
 * No reference counterpart (the reference is FP64 throughout). */
int wm_set_form(wm_ctx *ctx, int form);

/* ---- mpi_set__init (mpi_set.f90:19-49), single rank ---- */
int wm_mpi_set_init(int nygs, int nyge, int nproc, int *nys, int *nye,
                    int *nup, int *ndown, int *nrank, int *nsize);

/* ---- module initialisers ---- */
/* particle__init (particle.f90:18-45) */
int wm_particle_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs,
                     int nxge, int nygs, int nyge, int nys, int nye,
                     double delx, double delt, double c, const double *q,
                     const double *r);
/* field__init (field.f90:22-63); mnpr/ncomw/opsum/nerr are MPI handles that a
 * single device context does not use. */
int wm_field_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                  int nygs, int nyge, int nys, int nye, int mnpr, int ncomw,
                  int opsum, int nerr, double delx, double delt, double c,
                  const double *q, const double *r, double gfac);
/* sort__init (sort.f90:16-33) */
int wm_sort_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                 int nygs, int nyge, int nys, int nye);
/* boundary_periodic__init (boundary_periodic.f90:26-58) */
int wm_bcp_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                int nygs, int nyge, int nys, int nye, int nup, int ndown,
                int mnpi, int mnpr, int ncomw, int nerr, const int *nstat,
                int nstat_len, double delx, double delt, double c);
/* boundary_shock__init (proj/shock/boundary_shock.f90:27-59) and
 * boundary_reconnection__init (proj/reconnection/boundary_reconnection.f90:
 * 26-58): the same arguments; they select the wall family, whose host-array
 * calls take an active x range nxs..nxe inside nxgs..nxge.  A context takes
 * one family. */
int wm_bcs_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                int nygs, int nyge, int nys, int nye, int nup, int ndown,
                int mnpi, int mnpr, int ncomw, int nerr, const int *nstat,
                int nstat_len, double delx, double delt, double c);
int wm_bcr_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                int nygs, int nyge, int nys, int nye, int nup, int ndown,
                int mnpi, int mnpr, int ncomw, int nerr, const int *nstat,
                int nstat_len, double delx, double delt, double c);

/* ---- host-array entry points (reference call semantics) ---- */
/* particle__solv(gp,up,uf,cumcnt,nxs,nxe) (particle.f90:48-177) */
int wm_particle_solv(wm_ctx *ctx, double *gp, const double *up,
                     const double *uf, const int *cumcnt, int nxs, int nxe);
/* field__fdtd_i(uf,up,gp,cumcnt,nxs,nxe,bc_dfield,bc_curre,bc_phi)
 * (field.f90:66-186).  The three boundary callbacks of the reference are
 * selected by bc_family (WM_BC_*; negative: the family whose __init was
 * called) and run on the device.  The wall families solve with the
 * reference's plain CG. */
int wm_field_fdtd_i(wm_ctx *ctx, double *uf, const double *up,
                    const double *gp, const int *cumcnt, int nxs, int nxe,
                    int bc_family);
/* boundary_periodic__particle_x(up,np2) (boundary_periodic.f90:61-96) */
int wm_bcp_particle_x(wm_ctx *ctx, double *up, const int *np2);
/* boundary_periodic__particle_y(up,np2) (boundary_periodic.f90:99-248) */
int wm_bcp_particle_y(wm_ctx *ctx, double *up, int *np2);
/* sort__bucket(gp,up,cumcnt,np2,nxs,nxe) (sort.f90:36-82): gp is the output */
int wm_sort_bucket(wm_ctx *ctx, double *gp, const double *up, int *cumcnt,
                   const int *np2, int nxs, int nxe);
/* boundary_periodic__dfield(df,nxs,nxe,nys,nye,nxgs,nxge)
 * (boundary_periodic.f90:251-354) */
int wm_bcp_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye,
                  int nxgs, int nxge);
/* boundary_periodic__curre(uj,nxs,nxe,nys,nye,nxgs,nxge)
 * (boundary_periodic.f90:357-508) */
int wm_bcp_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye,
                 int nxgs, int nxge);
/* boundary_periodic__phi(phi,nxs,nxe,nys,nye,l) (boundary_periodic.f90:511-568) */
int wm_bcp_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye,
               int l);

/* ---- the shock family (proj/shock/boundary_shock.f90) ----
 * particle_x (:62-100, the reference's signature is (up,np2,nxs,nxe)):
 * mirrors about (nxs+1) delx and (nxe-1) delx with the momentum reversed.
 * injection (:255-297): the left mirror, and beyond xend = nxe delx +
 * u0/sqrt(1+u0^2/c^2) delt x mirrors about xend and u_x about u0.
 * particle_y (:103-252): the periodic one.  dfield (:300-407), curre
 * (:410-548), phi (:551-625): y periodic; a conducting wall at nxs (Bx, Ey,
 * Ez odd about it, By, Bz, Ex even about column nxs), zero beyond nxe, no x
 * fold of the current.  mom (:628-693): x ghosts fold into their own edge. */
int wm_bcs_particle_x(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe);
int wm_bcs_injection(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe,
                     double u0);
int wm_bcs_particle_y(wm_ctx *ctx, double *up, int *np2);
int wm_bcs_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye,
                  int nxgs, int nxge);
int wm_bcs_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye,
                 int nxgs, int nxge);
int wm_bcs_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye,
               int l);
int wm_bcs_mom(wm_ctx *ctx, double *mom);
/* ---- the reconnection family (proj/reconnection/boundary_reconnection.f90)
 * particle_x (:61-99), particle_y (:102-251), dfield (:254-361; the right
 * wall mirrors the left one about nxe - 1/2 and rewrites column nxe of Bx,
 * Ey, Ez), curre (:364-502), phi (:505-579), mom (:582-647). */
int wm_bcr_particle_x(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe);
int wm_bcr_particle_y(wm_ctx *ctx, double *up, int *np2);
int wm_bcr_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye,
                  int nxgs, int nxge);
int wm_bcr_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye,
                 int nxgs, int nxge);
int wm_bcr_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye,
               int l);
int wm_bcr_mom(wm_ctx *ctx, double *mom);
/* Device-resident runs of a wall family: the active range nxs..nxe and the
 * shock's u0, the arguments its loop passes every step
 * (proj/shock/app.f90:112-117).  No reference counterpart (the reference
 * passes them per call).  The wall families run in the FP64 form, with the
 * reference's plain CG. */
int wm_set_walls(wm_ctx *ctx, int nxs, int nxe, double u0);

/* ---- device-resident state and step ---- */
/* Ingest the reference's state (as after proj/weibel/app.f90:362 gp = up):
 * up (np2 active per row, sorted with cumcnt) and uf. */
int wm_state_upload(wm_ctx *ctx, const double *up, const int *np2,
                    const int *cumcnt, const double *uf);
/* Materialise the device state into the reference layout. */
int wm_state_download(wm_ctx *ctx, double *up, int *np2, int *cumcnt,
                      double *uf);
/* Synthetic counter-streaming pair plasma generated on the device:
 * ppc particles per cell per species, positions uniform in each row,
 * u = (0, 0, +/-udrift) + N(0, vth) per component, co-located species,
 * fields zero.  Same formula as wumingpic2d_amd.weibel.init_state. */
int wm_state_init_weibel(wm_ctx *ctx, uint64_t seed, int ppc, double vth,
                         double udrift);
/* nsteps steps of the reference loop body proj/weibel/app.f90:102-106:
 * particle__solv, field__fdtd_i, particle_x, particle_y, sort__bucket. */
int wm_step(wm_ctx *ctx, int nsteps);
/* energy_history sums (proj/weibel/app.f90:479-545) of the resident state:
 * out[0..nsp-1] kinetic per species, out[nsp] E, out[nsp+1] B. */
int wm_energy(wm_ctx *ctx, double *out);
/* CG iterations per component (l = 1..3) of the last field solve. */
int wm_cg_iters(wm_ctx *ctx, int *iters3);
/* GPU time (ms) of each phase of the last wm_step call, summed over its
 * steps, from HIP events on the context's stream: [0] cell-centred fields,
 * [1] push+deposit, [2] current gather+fold, [3] field solve, [4] particle
 * boundaries, [5] sort.  In the FP32 form the push places the last push's
 * movers (the particle boundaries and the sort reduce to that); [4] is the
 * bins' fill check beside the solve and [5] the re-layouts of rows whose
 * bins would overflow.  [6] and [7] hold the last wm_moments call's fused
 * moment pass and its fold.  Returns the number written. */
int wm_phase_times(wm_ctx *ctx, double *ms, int n);
/* Total active particles (sum of np2). */
int wm_particle_count(wm_ctx *ctx, long long *count);
/* The active particles' IDs (up(6,...) as 64-bit integers) summed and
 * xor-ed with 64-bit wrap, computed on the device: sum_xor[0], sum_xor[1].
 * A property of the particle set that no step may change, for states too
 * large to download.  No reference counterpart. */
int wm_id_checksum(wm_ctx *ctx, unsigned long long *sum_xor);
/* FP32 form: the particle store's capacities and their use since the state
 * was made (no reference counterpart; the reference stops when a row
 * overflows np, boundary_periodic.f90:228-236).  out[0] most particles in
 * one cell bin, [1] most particles leaving one bin in a push, [2] most
 * movers of one push workgroup, [3] most mover records of one row in a
 * push, [4] most entries of one row's spill of the push's mover lists, [5]
 * blocks of 64 cells whose bins were laid out anew (a bin would have
 * overflowed), [6] slots per (species, row) segment, [7] mover records per
 * row, [8] movers one push workgroup lists in LDS, [9] spill entries per
 * row, [10] bins found full, [11] segments laid out anew as a whole (a
 * block's slots could not hold its bins).  Returns the number written (at
 * most n, at most 12). */
int wm_capacity_stats(wm_ctx *ctx, long long *out, int n);

/* ---- moment diagnostics of an output step (proj/weibel/app.f90:119-121)
 *      mom (7, nxgs-1:nxge+1, nys-1:nye+1, nsp) real(8): N, Vx, Vy, Vz, Txx,
 *      Tyy, Tzz per node and species ---- */
/* mom_calc__init (mom_calc.f90:18-45); delt is the run's step, the module
 * advances u over delt/2 (mom_calc.f90:34). */
int wm_mom_calc_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs,
                     int nxge, int nygs, int nyge, int nys, int nye,
                     double delx, double delt, double c, const double *q,
                     const double *r);
/* mom_calc__accl(gp,up,uf,cumcnt,nxs,nxe) (mom_calc.f90:48-164): gp(1:5) of
 * the particles in cumcnt; gp(6) is not written. */
int wm_mom_calc_accl(wm_ctx *ctx, double *gp, const double *up,
                     const double *uf, const int *cumcnt, int nxs, int nxe);
/* mom_calc__nvt(mom,up,np2) (mom_calc.f90:167-249).  Positions must lie in
 * the cells nxgs..nxge of their rows (as after sort__bucket). */
int wm_mom_calc_nvt(wm_ctx *ctx, double *mom, const double *up,
                    const int *np2);
/* boundary_periodic__mom(mom) (boundary_periodic.f90:571-636) */
int wm_bcp_mom(wm_ctx *ctx, double *mom);
/* The three calls above on the device-resident state, fused (u over delt/2
 * and the moments in one pass over the particles, then the fold); only mom
 * is copied to the host.  Needs mom_calc__init. */
int wm_moments(wm_ctx *ctx, double *mom);

/* ---- restart files (common/paraio.f90): a JSON file of attributes and
 *      dataset offsets beside a raw file, for one process (nproc = 1) ---- */
/* paraio__init(ndim,np,nsp,nxgs,nxge,nygs,nyge,nys,nye,nproc,nrank,delx,
 * delt,c,q,r,dir) (paraio.f90:34-76); the grid must be the modules'. */
int wm_paraio_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge,
                   int nygs, int nyge, int nys, int nye, int nproc, int nrank,
                   double delx, double delt, double c, const double *q,
                   const double *r, const char *dir);
/* paraio__output(up,uf,np2,nxs,nxe,it,filename) (paraio.f90:93-228): writes
 * dir//filename//'.json' and '.raw' from host arrays. */
int wm_paraio_output(wm_ctx *ctx, const double *up, const double *uf,
                     const int *np2, int nxs, int nxe, int it,
                     const char *filename);
/* paraio__input(up,uf,np2,nxs,nxe,it,filename) (paraio.f90:233-426): reads
 * them back into host arrays, with the reference's checks and stops. */
int wm_paraio_input(wm_ctx *ctx, double *up, double *uf, int *np2, int *nxs,
                    int *nxe, int *it, const char *filename);
/* The same files from the device-resident state, row by row (no host copy
 * of the whole state; up's inactive slots are left holes of the file), and
 * into it: the rows as io__input reads them, sorted by cell as the app's
 * sort__bucket after io__input leaves them (proj/weibel/app.f90:351-352),
 * with the CG warm start cleared as in a fresh process. */
int wm_paraio_output_state(wm_ctx *ctx, int it, const char *filename);
int wm_paraio_input_state(wm_ctx *ctx, int *it, const char *filename);
/* The run's other outputs in the same format: paraio__param(n0,wpe,wpi,wge,
 * wgi,vti,vte,filename) (paraio.f90:431-555); paraio__mom(mom,uf,it)
 * (:557-713, divides mom's velocities and temperatures by the density in
 * place, as the reference does); paraio__ptcl(up,uf,np2,it) and
 * paraio__orb(up,uf,np2,it) (:715-863: every active particle, or those with
 * a positive ID).  File names are filename, or it followed by _mom, _ptcl or
 * _orb, under the directory of wm_paraio_init. */
int wm_paraio_param(wm_ctx *ctx, int n0, double wpe, double wpi, double wge,
                    double wgi, double vti, double vte, const char *filename);
int wm_paraio_mom(wm_ctx *ctx, double *mom, const double *uf, int it);
int wm_paraio_ptcl(wm_ctx *ctx, const double *up, const double *uf,
                   const int *np2, int it);
int wm_paraio_orb(wm_ctx *ctx, const double *up, const double *uf,
                  const int *np2, int it);

/* ---- learned preconditioner of the implicit solve (no reference
 *      counterpart: the reference runs unpreconditioned CG, field.f90:319-461).
 *      A four-level linear convolutional net M(r) ~ (f4 I - N4)^-1 r with BF16
 *      MFMA convolutions, trained on the device on (residual, correction)
 *      pairs that the plain solver records.  Needs nx, ny multiples of 8 and
 *      field__init (the operator depends on delx, delt, c, gfac). ---- */
/* WM_SOLVER_FCG: flexible CG with M, stopping at |r| <= err |b|, followed
 * by the plain solver's own residual test (and plain iterations from the
 * result if that test fails or a component went bad).  Needs parameters. */
int wm_set_solver(wm_ctx *ctx, int solver);
/* Record pairs: from now on each plain solve stores (r_k, x_final - x_k)
 * of its first per_solve iterations (all three components, one slot each)
 * in a ring of nslot slots (nslot <= 64). */
int wm_pc_collect(wm_ctx *ctx, int nslot, int per_solve);
/* Number of recorded pairs (slots filled). */
int wm_pc_pair_count(wm_ctx *ctx);
/* Default initial parameters from seed (also done by the first wm_pc_train). */
int wm_pc_init(wm_ctx *ctx, uint64_t seed);
/* Adam on the recorded pairs for `steps` steps, learning rate decaying
 * from lr by a cosine; loss2 (may be NULL) gets the first and the mean of
 * the last ten losses, L = mean((M(r)/s - e/s)^2), s = rms(r). */
int wm_pc_train(wm_ctx *ctx, int steps, double lr, uint64_t seed, double *loss2);
/* Number of FP32 parameters. */
int wm_pc_num_params(void);
/* Copy the FP32 parameters out (set = 0) or in (set = 1, resets Adam). */
int wm_pc_params(wm_ctx *ctx, float *theta, int n, int set);
/* z = M(r) for three host images r[3][ny][nx] (x fastest). */
int wm_pc_apply(wm_ctx *ctx, const double *r, double *z);
/* z = M(r) like wm_pc_apply, for device images r, z [3][ny][nx] (x fastest)
 * of any size with nx, ny positive multiples of 8 and nx * ny <= 262144,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  The activations live in a static device array of the
 * library (about 27 MB, held on each GPU that calls this until the process
 * ends), shared by every call on a GPU: calls that overlap on one GPU race.
 * The context supplies the parameters (set or trained beforehand); the call
 * returns when z is written.  Larger images return WM_ERR_ARG. */
int wm_pc_apply_noalloc(wm_ctx *ctx, const double *r, double *z, int nx,
                        int ny);
/* Loss and dL/dtheta for host images r, e [3][ny][nx]. */
int wm_pc_grad(wm_ctx *ctx, const double *r, const double *e, float *grad,
               double *loss);
/* Iterations of the last solve per component: flexible CG, then plain. */
int wm_solve_stats(wm_ctx *ctx, int *fcg3, int *plain3);

#ifdef __cplusplus
}
#endif
#endif
