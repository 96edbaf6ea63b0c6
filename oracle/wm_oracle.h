/*
 * wm_oracle.h -- CPU restatement (FP64, single thread) of the WumingPIC2D
 * per-step hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * that the HIP path is compared against.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Every array uses the reference's own column-major layout and index bounds
 * (Fortran order, first index fastest):
 *   up, gp : (ndim, np, nys:nye, nsp)        particle.f90:52-54
 *   uf, df : (6, nxgs-2:nxge+2, nys-2:nye+2)  particle.f90:53, field.f90:95
 *   uj     : (3, nxgs-2:nxge+2, nys-2:nye+2)  field.f90:195
 *   gkl    : (3, nxgs:nxge, nys:nye)          field.f90:107
 *   cumcnt : (nxgs:nxge+1, nys:nye, nsp)      particle.f90:51
 *   np2    : (nys:nye, nsp)                   boundary_periodic.f90:65
 *   phi    : (nxs-1:nxe+1, nys-1:nye+1)       boundary_periodic.f90:514
 * Single MPI rank: nys = nygs, nye = nyge, nxs = nxgs, nxe = nxge
 * (mpi_set.f90:122-132 with nsize = 1).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_particle_solv   bit-exact vs the amdflang build of common/particle.f90
 *                      (oracle/_ref) and committed golden vectors; closed-form
 *                      Boris gyration / E-acceleration; TSC linear reproduction.
 *   or_sort_bucket     bit-exact vs the amdflang build of common/sort.f90 and
 *                      golden vectors; brute-force stable argsort.
 *   or_ele_cur         discrete charge continuity (rho from the shape at old and
 *                      new positions), sum Jx = q dx/dt * displacement, Jz of a
 *                      resting particle, and Jz of moving (also cell-crossing)
 *                      particles against the time integral of the linearly
 *                      interpolated TSC shapes, by Gauss-Legendre quadrature
 *                      (test_oracle.py, test_oracle_field.py).
 *   or_cgm             converged solution solves the 5-point Helmholtz system;
 *                      Fourier-mode exact solution; eigenvector RHS takes exactly
 *                      two iterations (stop test on the pre-update residual);
 *                      |b| < 1e-6 skips the loop (squared initial sum_g).
 *   or_field_advance   (and or_field_fdtd_i) single Fourier modes of fields and
 *                      current against the implicit Maxwell scheme solved in
 *                      Fourier space; TE and TM vacuum eigenmodes against the
 *                      closed-form amplification factor of the theta scheme;
 *                      uniform J gives dE = -4 pi dt J and no dB; div B stays at
 *                      roundoff over 20 whole steps; Gauss law stays at
 *                      roundoff over 6 whole steps (test_oracle_field.py,
 *                      test_oracle.py).
 *   bc_* (halo/fold)   brute-force periodic index maps; fold conserves the sum.
 *   or_bc_particle_x   round-down wrap never lands on the upper bound.
 *   or_bc_particle_y   multiset and row membership invariants; hole-filling
 *                      ORDER is parity unpinned (no reference build exists for
 *                      boundary_periodic.f90; the order follows a reading of
 *                      boundary_periodic.f90:143-236 only).
 *   or_energy          exact sums of closed-form states.
 *   or_mom_calc_accl   bit-exact vs the amdflang build of common/mom_calc.f90;
 *   or_mom_calc_nvt    accl = particle__solv's velocities at delt/2; zero
 *                      fields leave u unchanged; a single particle's four node
 *                      weights (truncating int() below x = delx/2); partition
 *                      of unity of N and sum Vx = sum u gamma^-1.
 *   or_bc_mom          every ghost entry lands on its periodic image once.
 */
#ifndef WM_ORACLE_H
#define WM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXSP 8

typedef struct {
  int ndim, np, nsp;
  int nxgs, nxge, nygs, nyge, nys, nye;
  double delx, delt, c, gfac;
  double q[OR_MAXSP], r[OR_MAXSP];
  /* The active x range [nxs, nxe] the reference passes to particle__solv,
   * field__fdtd_i, sort__bucket (nxe = 0: the whole grid), the boundary
   * family (OR_BC_*), and the shock's upstream four-velocity u0
   * (proj/shock/app.f90:113). */
  int nxs, nxe, bc;
  double u0;
} or_params;

#define OR_BC_PERIODIC 0       /* common/boundary_periodic.f90 */
#define OR_BC_SHOCK 1          /* proj/shock/boundary_shock.f90 */
#define OR_BC_RECONNECTION 2   /* proj/reconnection/boundary_reconnection.f90 */

/* Error codes (mirror the reference's fatal stops). */
#define OR_OK 0
#define OR_ERR_CG_ITEMAX 1   /* field.f90:427-430 'stop at cgm after ite_max' */
#define OR_ERR_NP_OVER 2     /* boundary_periodic.f90:231-234 'memory over'   */
#define OR_ERR_RANGE 3       /* particle outside the range the reference indexes */
#define OR_ERR_ARG 4

/* Persistent SAVE state of field__fdtd_i (field.f90:96-119) and of
 * boundary_periodic__particle_y (boundary_periodic.f90:106-122). */
typedef struct or_state or_state;

or_state *or_state_new(const or_params *p);
void or_state_free(or_state *s);
/* Read-only views of the SAVE arrays (df, uj, gkl) for tests. */
double *or_state_df(or_state *s);
double *or_state_uj(or_state *s);
double *or_state_gkl(or_state *s);
/* CG iteration counts of the last field solve, per component l=1..3. */
void or_state_cg_iters(const or_state *s, int *iters3);
/* Override the CG relative tolerance (the reference fixes 1e-6). */
void or_state_set_cg_err(or_state *s, double err);

/* particle__solv (particle.f90:48-177) */
int or_particle_solv(const or_params *p, double *gp, const double *up,
                     const double *uf, const int *cumcnt);
/* ele_cur (field.f90:189-316) */
int or_ele_cur(const or_params *p, double *uj, const double *up,
               const double *gp, const int *cumcnt);
/* boundary_periodic__curre (boundary_periodic.f90:357-508) */
void or_bc_curre(const or_params *p, double *uj);
/* boundary_periodic__dfield (boundary_periodic.f90:251-354) */
void or_bc_dfield(const or_params *p, double *df);
/* boundary_periodic__phi (boundary_periodic.f90:511-568) */
void or_bc_phi(const or_params *p, double *phi);
/* bc_phi of component l (the wall families' ghost column depends on it) */
void or_bc_phi_l(const or_params *p, double *phi, int l);
/* boundary_shock__injection (boundary_shock.f90:255-298): the left wall and
 * the moving right edge xend = nxe delx + u0/sqrt(1+u0^2/c^2) delt. */
int or_bc_injection(const or_params *p, double *up, const int *np2, double u0);
/* RHS gkl of field__fdtd_i (field.f90:125-146) */
void or_field_rhs(const or_params *p, double *gkl, const double *uf,
                  const double *uj);
/* cgm (field.f90:319-461).  iters3 receives the iteration count per
 * component; returns OR_ERR_CG_ITEMAX when ite reaches ite_max. */
int or_cgm(const or_params *p, double *db, const double *gkl, double err,
           int *iters3);
/* field__fdtd_i (field.f90:66-186) using the periodic BC family:
 * ele_cur, bc_curre, then or_field_advance. */
int or_field_fdtd_i(const or_params *p, or_state *s, double *uf,
                    const double *up, const double *gp, const int *cumcnt);
/* field.f90:125-184 given the folded current uj (ghosts filled as after
 * boundary_periodic__curre): gkl, cgm, dE, both dfield halos, uf += df. */
int or_field_advance(const or_params *p, or_state *s, double *uf, const double *uj);
/* boundary_periodic__particle_x (boundary_periodic.f90:61-96) */
int or_bc_particle_x(const or_params *p, double *up, const int *np2);
/* boundary_periodic__particle_y (boundary_periodic.f90:99-248) */
int or_bc_particle_y(const or_params *p, or_state *s, double *up, int *np2);
/* sort__bucket (sort.f90:36-82) */
int or_sort_bucket(const or_params *p, double *gp, const double *up,
                   int *cumcnt, const int *np2);
/* mom_calc__accl (mom_calc.f90:48-164); p->delt is the run's step (the
 * module halves it, :34).  Writes gp(1:5) of the particles in cumcnt. */
int or_mom_calc_accl(const or_params *p, double *gp, const double *up,
                     const double *uf, const int *cumcnt);
/* mom_calc__nvt (mom_calc.f90:167-249): mom (7, nxgs-1:nxge+1, nys-1:nye+1,
 * nsp) = N, Vx, Vy, Vz, Txx, Tyy, Tzz (CIC sums) of up's np2 particles. */
int or_mom_calc_nvt(const or_params *p, double *mom, const double *up,
                    const int *np2);
/* boundary_periodic__mom (boundary_periodic.f90:571-636), one rank */
void or_bc_mom(const or_params *p, double *mom);
/* energy_history sums (proj/weibel/app.f90:479-545): out[0..nsp-1] kinetic
 * per species, out[nsp] E energy, out[nsp+1] B energy. */
void or_energy(const or_params *p, const double *up, const double *uf,
               const int *np2, double *out);
/* One step in the order of proj/weibel/app.f90:100-107:
 * push -> field -> bc_x -> bc_y -> sort (output back into up). */
int or_step(const or_params *p, or_state *s, double *up, double *gp,
            double *uf, int *cumcnt, int *np2);

/* Round-down (toward -inf) double addition, exact: the arithmetic the
 * reference obtains with ieee_set_rounding_mode(ieee_down)
 * (boundary_periodic.f90:74, 124). */
double or_add_rd(double a, double b);

/* Threads of the OpenMP loops (parity runs use 1). */
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
