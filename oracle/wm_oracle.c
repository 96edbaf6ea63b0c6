/*
 * wm_oracle.c -- FP64 CPU restatement of the WumingPIC2D hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, the smoke check and bench.py's CPU-baseline leg.
 *
 * Each function follows the reference routine it names line by line, in the
 * reference's evaluation order (left-to-right Fortran expression order, loop
 * nesting j -> i -> isp -> ii), so that a single-threaded build with FMA
 * contraction off reproduces the reference's rounding.  The OpenMP pragmas
 * mirror the reference's own !$OMP directives and are used only when the CPU
 * baseline is timed with several threads; parity runs use one thread.
 *
 * Citations are to files of WumingCode/WumingPIC2D.
 */
#include "wm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pi as the reference defines it: 4.0D0*atan(1.0D0) (field.f90:13). */
#define OR_PI 3.141592653589793115997963468544185161590576171875

/* ---------------------------------------------------------------- layout */
#define NXg(p) ((p)->nxge - (p)->nxgs + 1)
#define NYl(p) ((p)->nye - (p)->nys + 1)

static inline size_t iup(const or_params *p, int k, int ii, int j, int isp) {
  /* up(ndim, np, nys:nye, nsp), k/ii/isp 1-based */
  return (size_t)(k - 1) +
         (size_t)p->ndim * ((size_t)(ii - 1) +
                            (size_t)p->np * ((size_t)(j - p->nys) +
                                             (size_t)NYl(p) * (size_t)(isp - 1)));
}
static inline size_t iuf(const or_params *p, int c, int i, int j) {
  /* uf(6, nxgs-2:nxge+2, nys-2:nye+2) */
  return (size_t)(c - 1) +
         6u * ((size_t)(i - (p->nxgs - 2)) +
               (size_t)(NXg(p) + 4) * (size_t)(j - (p->nys - 2)));
}
static inline size_t iuj(const or_params *p, int c, int i, int j) {
  /* uj(3, nxgs-2:nxge+2, nys-2:nye+2) */
  return (size_t)(c - 1) +
         3u * ((size_t)(i - (p->nxgs - 2)) +
               (size_t)(NXg(p) + 4) * (size_t)(j - (p->nys - 2)));
}
static inline size_t icc(const or_params *p, int i, int j, int isp) {
  /* cumcnt(nxgs:nxge+1, nys:nye, nsp) */
  return (size_t)(i - p->nxgs) +
         (size_t)(NXg(p) + 1) *
             ((size_t)(j - p->nys) + (size_t)NYl(p) * (size_t)(isp - 1));
}
static inline size_t inp2(const or_params *p, int j, int isp) {
  return (size_t)(j - p->nys) + (size_t)NYl(p) * (size_t)(isp - 1);
}
static inline size_t igkl(const or_params *p, int c, int i, int j) {
  /* gkl(3, nxgs:nxge, nys:nye) */
  return (size_t)(c - 1) +
         3u * ((size_t)(i - p->nxgs) + (size_t)NXg(p) * (size_t)(j - p->nys));
}
static inline size_t iphi(const or_params *p, int i, int j) {
  /* phi(nxs-1:nxe+1, nys-1:nye+1) */
  return (size_t)(i - (p->nxgs - 1)) +
         (size_t)(NXg(p) + 2) * (size_t)(j - (p->nys - 1));
}

static size_t n_uf(const or_params *p) {
  return 6u * (size_t)(NXg(p) + 4) * (size_t)(NYl(p) + 4);
}
static size_t n_uj(const or_params *p) {
  return 3u * (size_t)(NXg(p) + 4) * (size_t)(NYl(p) + 4);
}
/* The active x range [nxs, nxe] (or_params; nxe = 0: nxgs..nxge). */
static inline int XS(const or_params *p) { return p->nxe > 0 ? p->nxs : p->nxgs; }
static inline int XE(const or_params *p) { return p->nxe > 0 ? p->nxe : p->nxge; }

/* ----------------------------------------------------------------- state */
struct or_state {
  or_params p;
  double *df, *gkl, *uj;   /* field.f90:98  SAVE, allocated once :105-119 */
  int *flag;               /* boundary_periodic.f90:110 flag(np, nys:nye)  */
  double *bff;             /* boundary_periodic.f90:111 bff_ptcl(ndim*np, nys-1:nye+1) */
  int iters[3];
  double cg_err;
};

or_state *or_state_new(const or_params *p) {
  or_state *s = (or_state *)calloc(1, sizeof(or_state));
  if (!s) return NULL;
  s->p = *p;
  s->df = (double *)calloc(n_uf(p), sizeof(double));
  s->uj = (double *)calloc(n_uj(p), sizeof(double));
  s->gkl = (double *)calloc(3u * (size_t)NXg(p) * (size_t)NYl(p), sizeof(double));
  s->flag = (int *)calloc((size_t)p->np * (size_t)NYl(p), sizeof(int));
  s->bff = (double *)calloc((size_t)p->ndim * (size_t)p->np * (size_t)(NYl(p) + 2),
                            sizeof(double));
  s->cg_err = 1e-6; /* field.f90:339 */
  if (!s->df || !s->uj || !s->gkl || !s->flag || !s->bff) {
    or_state_free(s);
    return NULL;
  }
  return s;
}
void or_state_free(or_state *s) {
  if (!s) return;
  free(s->df);
  free(s->uj);
  free(s->gkl);
  free(s->flag);
  free(s->bff);
  free(s);
}
double *or_state_df(or_state *s) { return s->df; }
double *or_state_uj(or_state *s) { return s->uj; }
double *or_state_gkl(or_state *s) { return s->gkl; }
void or_state_cg_iters(const or_state *s, int *it3) {
  it3[0] = s->iters[0];
  it3[1] = s->iters[1];
  it3[2] = s->iters[2];
}
void or_state_set_cg_err(or_state *s, double err) { s->cg_err = err; }

/* -------------------------------------------------------- round-down add */
double or_add_rd(double a, double b) {
  /* Knuth TwoSum gives the exact error of the rounded-to-nearest sum; a
   * negative error means the exact sum lies below s, so the value rounded
   * toward -inf is the next double below s. */
  volatile double s = a + b;
  double bb = s - a;
  double err = (a - (s - bb)) + (b - bb);
  if (err < 0.0) return nextafter(s, -INFINITY);
  return s;
}

/* ------------------------------------------------------- particle__solv */
int or_particle_solv(const or_params *p, double *gp, const double *up,
                     const double *uf, const int *cumcnt) {
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  const int nxt = nxe - nxs + 3, nyt = nye - nys + 3;
  const double d_delx = 1. / p->delx;           /* particle.f90:41 */
  const double delt = p->delt, c = p->c;
  double *tmp = (double *)malloc(sizeof(double) * 6u * (size_t)nxt * (size_t)nyt);
  if (!tmp) return OR_ERR_ARG;
#define TMP(k, i, j) tmp[(size_t)((k) - 1) + 6u * ((size_t)((i) - (nxs - 1)) + (size_t)nxt * (size_t)((j) - (nys - 1)))]
#define UF(k, i, j) uf[iuf(p, k, i, j)]

  /* fields at (i+1/2, j+1/2): particle.f90:68-81 */
#pragma omp parallel for
  for (int j = nys - 1; j <= nye + 1; ++j)
    for (int i = nxs - 1; i <= nxe + 1; ++i) {
      TMP(1, i, j) = 0.5 * (UF(1, i, j) + UF(1, i, j + 1));
      TMP(2, i, j) = 0.5 * (UF(2, i, j) + UF(2, i + 1, j));
      TMP(3, i, j) = 0.25 * (UF(3, i, j) + UF(3, i + 1, j) + UF(3, i, j + 1) +
                             UF(3, i + 1, j + 1));
      TMP(4, i, j) = 0.5 * (UF(4, i, j) + UF(4, i + 1, j));
      TMP(5, i, j) = 0.5 * (UF(5, i, j) + UF(5, i, j + 1));
      TMP(6, i, j) = UF(6, i, j);
    }

  /* particle loop: particle.f90:83-169 */
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (int j = nys; j <= nye; ++j)
    for (int i = nxs; i <= nxe; ++i)
      for (int isp = 1; isp <= p->nsp; ++isp) {
        const double q = p->q[isp - 1], r = p->r[isp - 1];
        const double fac1 = q / r * 0.5 * delt;  /* :90 */
        const double txxx = fac1 * fac1;         /* :91 */
        const double fac2 = q * delt / r;        /* :92 */
        const int i0 = cumcnt[icc(p, i, j, isp)], i1 = cumcnt[icc(p, i + 1, j, isp)];
        if (i0 < 0 || i1 > p->np) { bad = 1; continue; }   /* i1 < i0: an empty range */
        for (int ii = i0 + 1; ii <= i1; ++ii) {
          const double *u = up + iup(p, 1, ii, j, isp);
          double *g = gp + iup(p, 1, ii, j, isp);
          double sh[3][2], dh;
          /* second order shape function :96-105 */
          dh = u[0] * d_delx - 0.5 - i;
          sh[0][0] = 0.5 * (0.5 - dh) * (0.5 - dh);
          sh[1][0] = 0.75 - dh * dh;
          sh[2][0] = 0.5 * (0.5 + dh) * (0.5 + dh);
          dh = u[1] * d_delx - 0.5 - j;
          sh[0][1] = 0.5 * (0.5 - dh) * (0.5 - dh);
          sh[1][1] = 0.75 - dh * dh;
          sh[2][1] = 0.5 * (0.5 + dh) * (0.5 + dh);
          /* gather :107-129 */
          double f[6];
          for (int k = 1; k <= 6; ++k) {
            f[k - 1] =
                (TMP(k, i - 1, j - 1) * sh[0][0] + TMP(k, i, j - 1) * sh[1][0] +
                 TMP(k, i + 1, j - 1) * sh[2][0]) * sh[0][1] +
                (TMP(k, i - 1, j) * sh[0][0] + TMP(k, i, j) * sh[1][0] +
                 TMP(k, i + 1, j) * sh[2][0]) * sh[1][1] +
                (TMP(k, i - 1, j + 1) * sh[0][0] + TMP(k, i, j + 1) * sh[1][0] +
                 TMP(k, i + 1, j + 1) * sh[2][0]) * sh[2][1];
          }
          const double bpx = f[0], bpy = f[1], bpz = f[2];
          const double epx = f[3], epy = f[4], epz = f[5];
          /* accel. :132-134 */
          double uvm1 = u[2] + fac1 * epx;
          double uvm2 = u[3] + fac1 * epy;
          double uvm3 = u[4] + fac1 * epz;
          /* rotate :137-148 */
          double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
          const double igam = 1. / gam;
          const double fac1r = fac1 * igam;
          const double fac2r =
              fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
          const double uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
          const double uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
          const double uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
          uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
          uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
          uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
          /* accel. :151-153 */
          g[2] = uvm1 + fac1 * epx;
          g[3] = uvm2 + fac1 * epy;
          g[4] = uvm3 + fac1 * epz;
          /* move :156-161 */
          gam = 1. / sqrt(1.0 + (g[2] * g[2] + g[3] * g[3] + g[4] * g[4]) / (c * c));
          g[0] = u[0] + g[2] * delt * gam;
          g[1] = u[1] + g[3] * delt * gam;
        }
      }
  free(tmp);
#undef TMP
#undef UF
  if (bad) return OR_ERR_RANGE;
  /* ID copy over the whole capacity: particle.f90:171-175 */
  if (p->ndim == 6) {
    const size_t n = (size_t)p->np * (size_t)NYl(p) * (size_t)p->nsp;
#pragma omp parallel for
    for (size_t t = 0; t < n; ++t) gp[6 * t + 5] = up[6 * t + 5];
  }
  return OR_OK;
}

/* --------------------------------------------------------------- ele_cur */
/* Deposit of rows j0..j1 into uj (field.f90:212-313, one thread's share). */
static int ele_cur_rows(const or_params *p, double *uj, const double *up,
                        const double *gp, const int *cumcnt, int j0, int j1) {
  const int nxs = XS(p), nxe = XE(p);
  const double d_delx = 1. / p->delx, d_delt = 1. / p->delt;
  const double delx = p->delx, c = p->c;
  const double fac = 1.0 / 3.0; /* field.f90:198 */
  int bad = 0;
  for (int j = j0; j <= j1; ++j)
    for (int i = nxs; i <= nxe; ++i) {
      double pjx[5][5], pjy[5][5], pjz[5][5], pjtmp[5][5]; /* [ip+2][jp+2] */
      memset(pjx, 0, sizeof pjx);
      memset(pjy, 0, sizeof pjy);
      memset(pjz, 0, sizeof pjz);
      for (int isp = 1; isp <= p->nsp; ++isp) {
        const double q = p->q[isp - 1];
        const int i0 = cumcnt[icc(p, i, j, isp)], i1 = cumcnt[icc(p, i + 1, j, isp)];
        if (i0 < 0 || i1 > p->np) { bad = 1; continue; }   /* i1 < i0: an empty range */
        for (int ii = i0 + 1; ii <= i1; ++ii) {
          const double *u = up + iup(p, 1, ii, j, isp);
          const double *g = gp + iup(p, 1, ii, j, isp);
          double s0[5][2], ds[5][2], dh;
          /* :224-236 */
          dh = u[0] * d_delx - 0.5 - i;
          s0[0][0] = 0.0;
          s0[1][0] = 0.5 * (0.5 - dh) * (0.5 - dh);
          s0[2][0] = 0.75 - dh * dh;
          s0[3][0] = 0.5 * (0.5 + dh) * (0.5 + dh);
          s0[4][0] = 0.0;
          dh = u[1] * d_delx - 0.5 - j;
          s0[0][1] = 0.0;
          s0[1][1] = 0.5 * (0.5 - dh) * (0.5 - dh);
          s0[2][1] = 0.75 - dh * dh;
          s0[3][1] = 0.5 * (0.5 + dh) * (0.5 + dh);
          s0[4][1] = 0.0;
          /* :238-266 */
          for (int d = 0; d < 2; ++d) {
            const int base = d == 0 ? i : j;
            const int i2 = (int)(g[d] * d_delx);
            dh = g[d] * d_delx - 0.5 - i2;
            const int inc = i2 - base;
            if (inc < -1 || inc > 1) bad = 1;
            const double s1_1 = 0.5 * (0.5 - dh) * (0.5 - dh);
            const double s1_2 = 0.75 - dh * dh;
            const double s1_3 = 0.5 * (0.5 + dh) * (0.5 + dh);
            const int ainc = inc < 0 ? -inc : inc;
            const double smo_1 = -((double)(inc - ainc) * 0.5) + 0;
            const double smo_2 = (double)(-ainc + 1);
            const double smo_3 = (double)(inc + ainc) * 0.5 + 0;
            ds[0][d] = s1_1 * smo_1;
            ds[1][d] = s1_1 * smo_2 + s1_2 * smo_1;
            ds[2][d] = s1_2 * smo_2 + s1_3 * smo_1 + s1_1 * smo_3;
            ds[3][d] = s1_3 * smo_2 + s1_2 * smo_3;
            ds[4][d] = s1_3 * smo_3;
          }
          /* :268 */
          for (int k = 0; k < 5; ++k) {
            ds[k][0] = ds[k][0] - s0[k][0];
            ds[k][1] = ds[k][1] - s0[k][1];
          }
          /* :270-272 */
          const double gvz =
              g[4] / sqrt(1. + (g[2] * g[2] + g[3] * g[3] + g[4] * g[4]) / (c * c));
          /* Jx :274-281 */
          memset(pjtmp, 0, sizeof pjtmp);
          for (int jp = -2; jp <= 2; ++jp)
            for (int ip = -2; ip <= 1; ++ip)
              pjtmp[ip + 3][jp + 2] =
                  pjtmp[ip + 2][jp + 2] -
                  q * delx * d_delt * ds[ip + 2][0] * (s0[jp + 2][1] + 0.5 * ds[jp + 2][1]);
          for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b) pjx[a][b] = pjx[a][b] + pjtmp[a][b];
          /* Jy :283-290 */
          memset(pjtmp, 0, sizeof pjtmp);
          for (int jp = -2; jp <= 1; ++jp)
            for (int ip = -2; ip <= 2; ++ip)
              pjtmp[ip + 2][jp + 3] =
                  pjtmp[ip + 2][jp + 2] -
                  q * delx * d_delt * ds[jp + 2][1] * (s0[ip + 2][0] + 0.5 * ds[ip + 2][0]);
          for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b) pjy[a][b] = pjy[a][b] + pjtmp[a][b];
          /* Jz :292-298 */
          for (int jp = -2; jp <= 2; ++jp)
            for (int ip = -2; ip <= 2; ++ip)
              pjz[ip + 2][jp + 2] =
                  pjz[ip + 2][jp + 2] +
                  q * gvz *
                      (s0[ip + 2][0] * s0[jp + 2][1] + 0.5 * ds[ip + 2][0] * s0[jp + 2][1] +
                       0.5 * s0[ip + 2][0] * ds[jp + 2][1] +
                       fac * ds[ip + 2][0] * ds[jp + 2][1]);
        }
      }
      /* scatter :304-310 */
      for (int jp = -2; jp <= 2; ++jp)
        for (int ip = -2; ip <= 2; ++ip) {
          uj[iuj(p, 1, i + ip, j + jp)] += pjx[ip + 2][jp + 2];
          uj[iuj(p, 2, i + ip, j + jp)] += pjy[ip + 2][jp + 2];
          uj[iuj(p, 3, i + ip, j + jp)] += pjz[ip + 2][jp + 2];
        }
    }
  return bad;
}

int or_ele_cur(const or_params *p, double *uj, const double *up,
               const double *gp, const int *cumcnt) {
  const size_t nuj = n_uj(p);
  /* field.f90:203-205 */
  memset(uj, 0, sizeof(double) * nuj);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  const int ny = NYl(p);
  if (nt > ny) nt = ny;
  if (nt <= 1)
    return ele_cur_rows(p, uj, up, gp, cumcnt, p->nys, p->nye) ? OR_ERR_RANGE : OR_OK;
  /* REDUCTION(+:uj) of field.f90:211: one private copy per thread, summed
   * into uj in thread order afterwards. */
  double *priv = (double *)calloc((size_t)nt * nuj, sizeof(double));
  if (!priv) return OR_ERR_ARG;
  int bad = 0;
#pragma omp parallel num_threads(nt) reduction(| : bad)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    const int j0 = p->nys + (int)((long)ny * t / nt);
    const int j1 = p->nys + (int)((long)ny * (t + 1) / nt) - 1;
    bad |= ele_cur_rows(p, priv + (size_t)t * nuj, up, gp, cumcnt, j0, j1);
  }
#pragma omp parallel for
  for (size_t k = 0; k < nuj; ++k) {
    double a = uj[k];
    for (int t = 0; t < nt; ++t) a += priv[(size_t)t * nuj + k];
    uj[k] = a;
  }
  free(priv);
  return bad ? OR_ERR_RANGE : OR_OK;
}

/* ------------------------------------------------- boundary_periodic__* */
void or_bc_curre(const or_params *p, double *uj) {
  /* boundary_periodic.f90:357-508 with nup = ndown = self. */
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  const int nb = nxe - nxs + 5;
  double *snd = (double *)malloc(sizeof(double) * 6u * (size_t)nb);
  double *rcv = (double *)malloc(sizeof(double) * 6u * (size_t)nb);
#define UJ(k, i, j) uj[iuj(p, k, i, j)]
  /* send rows nys-2, nys-1 to rank-1 :370-384 */
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      snd[ii + k - 1] = UJ(k, i, nys - 2);
      snd[ii + 3 + k - 1] = UJ(k, i, nys - 1);
    }
  }
  memcpy(rcv, snd, sizeof(double) * 6u * (size_t)nb);
  /* :388-398 and :401-411 (the second loop reads rows nye+1, nye+2) */
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      UJ(k, i, nye - 1) = UJ(k, i, nye - 1) + rcv[ii + k - 1];
      UJ(k, i, nye) = UJ(k, i, nye) + rcv[ii + 3 + k - 1];
    }
  }
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      snd[ii + k - 1] = UJ(k, i, nye + 1);
      snd[ii + 3 + k - 1] = UJ(k, i, nye + 2);
    }
  }
  memcpy(rcv, snd, sizeof(double) * 6u * (size_t)nb);
  /* :419-429 */
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      UJ(k, i, nys) = UJ(k, i, nys) + rcv[ii + k - 1];
      UJ(k, i, nys + 1) = UJ(k, i, nys + 1) + rcv[ii + 3 + k - 1];
    }
  }
  /* copy back into the ghost rows :433-493 */
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      snd[ii + k - 1] = UJ(k, i, nys);
      snd[ii + 3 + k - 1] = UJ(k, i, nys + 1);
    }
  }
  memcpy(rcv, snd, sizeof(double) * 6u * (size_t)nb);
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      UJ(k, i, nye + 1) = rcv[ii + k - 1];
      UJ(k, i, nye + 2) = rcv[ii + 3 + k - 1];
    }
  }
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      snd[ii + k - 1] = UJ(k, i, nye - 1);
      snd[ii + 3 + k - 1] = UJ(k, i, nye);
    }
  }
  memcpy(rcv, snd, sizeof(double) * 6u * (size_t)nb);
  for (int i = nxs - 2; i <= nxe + 2; ++i) {
    const int ii = 6 * (i - (nxs - 2));
    for (int k = 1; k <= 3; ++k) {
      UJ(k, i, nys - 2) = rcv[ii + k - 1];
      UJ(k, i, nys - 1) = rcv[ii + 3 + k - 1];
    }
  }
  /* columns :495-506 (whole-array statements, executed in order); the wall
   * families fold nothing in x (boundary_shock.f90:410-548,
   * boundary_reconnection.f90:364-502 end after the rows) */
  if (p->bc != OR_BC_PERIODIC) {
    free(snd);
    free(rcv);
    return;
  }
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxe - 1, j) = UJ(k, nxe - 1, j) + UJ(k, nxs - 2, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxe, j) = UJ(k, nxe, j) + UJ(k, nxs - 1, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxs, j) = UJ(k, nxs, j) + UJ(k, nxe + 1, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxs + 1, j) = UJ(k, nxs + 1, j) + UJ(k, nxe + 2, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxs - 2, j) = UJ(k, nxe - 1, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxs - 1, j) = UJ(k, nxe, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxe + 1, j) = UJ(k, nxs, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 3; ++k) UJ(k, nxe + 2, j) = UJ(k, nxs + 1, j);
#undef UJ
  free(snd);
  free(rcv);
}

void or_bc_dfield(const or_params *p, double *df) {
  /* boundary_periodic.f90:251-354 with nup = ndown = self. */
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
#define DF(k, i, j) df[iuf(p, k, i, j)]
  /* rows nys, nys+1 -> nye+1, nye+2 :263-302 */
  for (int i = nxs; i <= nxe; ++i)
    for (int k = 1; k <= 6; ++k) {
      DF(k, i, nye + 1) = DF(k, i, nys);
      DF(k, i, nye + 2) = DF(k, i, nys + 1);
    }
  /* rows nye-1, nye -> nys-2, nys-1 :305-344 (read after the first copy) */
  for (int i = nxs; i <= nxe; ++i)
    for (int k = 1; k <= 6; ++k) {
      const double a = DF(k, i, nye - 1), b = DF(k, i, nye);
      DF(k, i, nys - 2) = a;
      DF(k, i, nys - 1) = b;
    }
  if (p->bc == OR_BC_SHOCK) {
    /* conducting wall at the left, zero at the right edge
     * (boundary_shock.f90:396-405) */
    for (int j = nys - 2; j <= nye + 2; ++j) {
      DF(1, nxs - 1, j) = -DF(1, nxs, j);
      for (int k = 2; k <= 4; ++k) DF(k, nxs - 1, j) = DF(k, nxs + 1, j);
      for (int k = 5; k <= 6; ++k) DF(k, nxs - 1, j) = -DF(k, nxs, j);
      for (int k = 1; k <= 6; ++k) DF(k, nxe + 1, j) = 0.0;
    }
    return;
  }
  if (p->bc == OR_BC_RECONNECTION) {
    /* conducting walls on both sides; the right one rewrites the last
     * interior column of Bx, Ey, Ez (boundary_reconnection.f90:350-359) */
    for (int j = nys - 2; j <= nye + 2; ++j) {
      DF(1, nxs - 1, j) = -DF(1, nxs, j);
      for (int k = 2; k <= 4; ++k) DF(k, nxs - 1, j) = DF(k, nxs + 1, j);
      for (int k = 5; k <= 6; ++k) DF(k, nxs - 1, j) = -DF(k, nxs, j);
      DF(1, nxe, j) = -DF(1, nxe - 1, j);
      for (int k = 2; k <= 4; ++k) DF(k, nxe + 1, j) = DF(k, nxe - 1, j);
      for (int k = 5; k <= 6; ++k) DF(k, nxe, j) = -DF(k, nxe - 1, j);
    }
    return;
  }
  /* columns :347-352 */
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 6; ++k) DF(k, nxs - 2, j) = DF(k, nxe - 1, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 6; ++k) DF(k, nxs - 1, j) = DF(k, nxe, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 6; ++k) DF(k, nxe + 1, j) = DF(k, nxs, j);
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int k = 1; k <= 6; ++k) DF(k, nxe + 2, j) = DF(k, nxs + 1, j);
#undef DF
}

void or_bc_phi_l(const or_params *p, double *phi, int l) {
  /* boundary_periodic.f90:511-568; the wall families' rows are the same
   * (boundary_shock.f90:551-602) and their columns depend on l (:603-623,
   * boundary_reconnection.f90:557-577): Bx odd about the wall, By and Bz
   * even about column nxs, zero beyond nxe */
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  for (int i = nxs; i <= nxe; ++i) phi[iphi(p, i, nye + 1)] = phi[iphi(p, i, nys)];
  for (int i = nxs; i <= nxe; ++i) phi[iphi(p, i, nys - 1)] = phi[iphi(p, i, nye)];
  if (p->bc != OR_BC_PERIODIC) {
    for (int j = nys - 1; j <= nye + 1; ++j) {
      phi[iphi(p, nxs - 1, j)] = l == 1 ? -phi[iphi(p, nxs, j)] : phi[iphi(p, nxs + 1, j)];
      phi[iphi(p, nxe + 1, j)] = 0.0;
    }
    return;
  }
  for (int j = nys - 1; j <= nye + 1; ++j) phi[iphi(p, nxs - 1, j)] = phi[iphi(p, nxe, j)];
  for (int j = nys - 1; j <= nye + 1; ++j) phi[iphi(p, nxe + 1, j)] = phi[iphi(p, nxs, j)];
}
void or_bc_phi(const or_params *p, double *phi) { or_bc_phi_l(p, phi, 1); }

/* ------------------------------------------------------------ field RHS */
static void field_consts(const or_params *p, double *f1, double *f2, double *f3,
                         double *f4, double *f5) {
  /* field.f90:53-57 */
  const double c = p->c, delt = p->delt, delx = p->delx, gfac = p->gfac;
  *f1 = c * delt / delx;
  *f2 = gfac * *f1 * *f1;
  *f3 = 4.0 * OR_PI * delx / c;
  const double t = delx / (c * delt * gfac);
  *f4 = 4.0 + t * t;
  *f5 = t * t;
}

void or_field_rhs(const or_params *p, double *gkl, const double *uf,
                  const double *uj) {
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  double f1, f2, f3, f4, f5;
  field_consts(p, &f1, &f2, &f3, &f4, &f5);
  (void)f4;
  (void)f5;
#define UF(k, i, j) uf[iuf(p, k, i, j)]
#define UJ(k, i, j) uj[iuj(p, k, i, j)]
  /* field.f90:125-146 */
#pragma omp parallel for
  for (int j = nys; j <= nye; ++j)
    for (int i = nxs; i <= nxe; ++i) {
      gkl[igkl(p, 1, i, j)] =
          f2 * (UF(1, i, j - 1) + UF(1, i - 1, j) - 4. * UF(1, i, j) + UF(1, i + 1, j) +
                UF(1, i, j + 1) + f3 * (-UJ(3, i, j - 1) + UJ(3, i, j))) -
          f1 * (-UF(6, i, j - 1) + UF(6, i, j));
      gkl[igkl(p, 2, i, j)] =
          f2 * (UF(2, i, j - 1) + UF(2, i - 1, j) - 4. * UF(2, i, j) + UF(2, i + 1, j) +
                UF(2, i, j + 1) - f3 * (-UJ(3, i - 1, j) + UJ(3, i, j))) +
          f1 * (-UF(6, i - 1, j) + UF(6, i, j));
      gkl[igkl(p, 3, i, j)] =
          f2 * (UF(3, i, j - 1) + UF(3, i - 1, j) - 4. * UF(3, i, j) + UF(3, i + 1, j) +
                UF(3, i, j + 1) +
                f3 * (-UJ(2, i - 1, j) + UJ(2, i, j) + UJ(1, i, j - 1) - UJ(1, i, j))) -
          f1 * (-UF(5, i - 1, j) + UF(5, i, j) + UF(4, i, j - 1) - UF(4, i, j));
    }
#undef UF
#undef UJ
}

/* ------------------------------------------------------------------ cgm */
int or_cgm(const or_params *p, double *db, const double *gkl, double err,
           int *iters3) {
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  const int ite_max = 100; /* field.f90:337 */
  double f1, f2, f3, f4, f5;
  field_consts(p, &f1, &f2, &f3, &f4, &f5);
  /* phi and p span nxgs-1..nxge+1 (iphi); the solve uses nxs-1..nxe+1 */
  const size_t nphi = (size_t)(NXg(p) + 2) * (size_t)(nye - nys + 3);
  const size_t nin = (size_t)(nxe - nxs + 1) * (size_t)(nye - nys + 1);
  double *phi = (double *)calloc(nphi, sizeof(double));
  double *pp = (double *)calloc(nphi, sizeof(double));
  double *r = (double *)calloc(nin, sizeof(double));
  double *b = (double *)calloc(nin, sizeof(double));
  double *ap = (double *)calloc(nin, sizeof(double));
  const int nxl = nxe - nxs + 1;
#define IN(i, j) ((size_t)((i) - nxs) + (size_t)nxl * (size_t)((j) - nys))
#define PH(a, i, j) a[iphi(p, i, j)]
  int status = OR_OK;
  for (int l = 1; l <= 3; ++l) {
    int ite = 0;
    double sum = 0.0;
    /* initial guess :350-360 */
    for (int j = nys; j <= nye; ++j)
      for (int i = nxs; i <= nxe; ++i) {
        PH(phi, i, j) = db[iuf(p, l, i, j)];
        b[IN(i, j)] = f5 * gkl[igkl(p, l, i, j)];
        sum = sum + b[IN(i, j)] * b[IN(i, j)];
      }
    double sum_g = sum;                /* :362 (single rank) */
    const double eps = sqrt(sum_g) * err; /* :364 */
    or_bc_phi_l(p, phi, l);            /* :367 */
    double sumr = 0.0;
    for (int j = nys; j <= nye; ++j)
      for (int i = nxs; i <= nxe; ++i) {
        r[IN(i, j)] = b[IN(i, j)] + PH(phi, i, j - 1) + PH(phi, i - 1, j) -
                      f4 * PH(phi, i, j) + PH(phi, i + 1, j) + PH(phi, i, j + 1);
        PH(pp, i, j) = r[IN(i, j)];
        sumr = sumr + r[IN(i, j)] * r[IN(i, j)];
      }
    double sumr_g = sumr;
    if (sqrt(sumr_g) > eps) {          /* :385 */
      while (sum_g > eps) {            /* :387 -- sum_g starts as |b|^2 */
        ite = ite + 1;
        or_bc_phi_l(p, pp, l);         /* :392 */
        sumr = 0.0;
        double sum2 = 0.0;
        for (int j = nys; j <= nye; ++j)
          for (int i = nxs; i <= nxe; ++i) {
            ap[IN(i, j)] = -PH(pp, i, j - 1) - PH(pp, i - 1, j) + f4 * PH(pp, i, j) -
                           PH(pp, i + 1, j) - PH(pp, i, j + 1);
            sumr = sumr + r[IN(i, j)] * r[IN(i, j)];
            sum2 = sum2 + PH(pp, i, j) * ap[IN(i, j)];
          }
        sumr_g = sumr;
        const double sum2_g = sum2;
        const double av = sumr_g / sum2_g; /* :415 */
        for (int j = nys; j <= nye; ++j)
          for (int i = nxs; i <= nxe; ++i) {
            PH(phi, i, j) = PH(phi, i, j) + av * PH(pp, i, j);
            r[IN(i, j)] = r[IN(i, j)] - av * ap[IN(i, j)];
          }
        sum_g = sqrt(sumr_g);          /* :426 -- the pre-update residual */
        if (ite >= ite_max) {          /* :427-430 */
          status = OR_ERR_CG_ITEMAX;
          break;
        }
        double sum1 = 0.0;
        for (int j = nys; j <= nye; ++j)
          for (int i = nxs; i <= nxe; ++i) sum1 = sum1 + r[IN(i, j)] * r[IN(i, j)];
        const double bv = sum1 / sumr_g; /* :442 */
        for (int j = nys; j <= nye; ++j)
          for (int i = nxs; i <= nxe; ++i)
            PH(pp, i, j) = r[IN(i, j)] + bv * PH(pp, i, j);
      }
    }
    if (iters3) iters3[l - 1] = ite;
    if (status != OR_OK) break;
    /* :455-457 */
    for (int j = nys; j <= nye; ++j)
      for (int i = nxs; i <= nxe; ++i) db[iuf(p, l, i, j)] = PH(phi, i, j);
  }
#undef IN
#undef PH
  free(phi);
  free(pp);
  free(r);
  free(b);
  free(ap);
  return status;
}

/* -------------------------------------------------------- field__fdtd_i */
int or_field_fdtd_i(const or_params *p, or_state *s, double *uf,
                    const double *up, const double *gp, const int *cumcnt) {
  int st = or_ele_cur(p, s->uj, up, gp, cumcnt);   /* :121 */
  if (st) return st;
  or_bc_curre(p, s->uj);                           /* :122 */
  return or_field_advance(p, s, uf, s->uj);        /* :125-184 */
}

int or_field_advance(const or_params *p, or_state *s, double *uf, const double *uj) {
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  double f1, f2, f3, f4, f5;
  field_consts(p, &f1, &f2, &f3, &f4, &f5);
  const double gfac = p->gfac, delt = p->delt;
  or_field_rhs(p, s->gkl, uf, uj);                 /* :125-146 */
  int st = or_cgm(p, s->df, s->gkl, s->cg_err, s->iters); /* :149 */
  if (st) return st;
  or_bc_dfield(p, s->df);                          /* :151 */
  double *df = s->df;
#define UF(k, i, j) uf[iuf(p, k, i, j)]
#define DF(k, i, j) df[iuf(p, k, i, j)]
#define UJ(k, i, j) uj[iuj(p, k, i, j)]
  /* field.f90:154-171 */
#pragma omp parallel for
  for (int j = nys; j <= nye; ++j)
    for (int i = nxs; i <= nxe; ++i) {
      DF(4, i, j) = f1 * (gfac * (-DF(3, i, j) + DF(3, i, j + 1)) +
                          (-UF(3, i, j) + UF(3, i, j + 1))) -
                    4. * OR_PI * delt * UJ(1, i, j);
      DF(5, i, j) = -f1 * (gfac * (-DF(3, i, j) + DF(3, i + 1, j)) +
                           (-UF(3, i, j) + UF(3, i + 1, j))) -
                    4. * OR_PI * delt * UJ(2, i, j);
      DF(6, i, j) = f1 * (gfac * (-DF(2, i, j) + DF(2, i + 1, j) + DF(1, i, j) -
                                  DF(1, i, j + 1)) +
                          (-UF(2, i, j) + UF(2, i + 1, j) + UF(1, i, j) - UF(1, i, j + 1))) -
                    4. * OR_PI * delt * UJ(3, i, j);
    }
  or_bc_dfield(p, df);                             /* :173 */
  /* :176-184 */
#pragma omp parallel for
  for (int j = nys - 2; j <= nye + 2; ++j)
    for (int i = nxs - 2; i <= nxe + 2; ++i)
      for (int k = 1; k <= 6; ++k) UF(k, i, j) = UF(k, i, j) + DF(k, i, j);
#undef UF
#undef DF
#undef UJ
  return OR_OK;
}

/* ---------------------------------------------- particle boundary: x, y */
/* The wall families' particle_x (boundary_shock.f90:62-100 and
 * boundary_reconnection.f90:61-99, the same statements): a particle in a
 * cell below nxs+1 is mirrored about x = (nxs+1) delx, one in a cell at or
 * above nxe-1 about x = (nxe-1) delx, and its momentum reversed. */
static void wall_x(const or_params *p, double *up, const int *np2) {
  const int nxs = XS(p), nxe = XE(p);
  const double delx = p->delx;
  for (int isp = 1; isp <= p->nsp; ++isp)
    for (int j = p->nys; j <= p->nye; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        double *u = up + iup(p, 1, ii, j, isp);
        const int ipos = (int)(u[0] / delx);
        double w;
        if (ipos < nxs + 1)
          w = 2. * (nxs + 1) * delx;
        else if (ipos >= nxe - 1)
          w = 2. * (nxe - 1) * delx;
        else
          continue;
        u[0] = w - u[0];
        u[2] = -u[2];
        u[3] = -u[3];
        u[4] = -u[4];
      }
}

int or_bc_injection(const or_params *p, double *up, const int *np2, double u0) {
  /* boundary_shock.f90:255-298: the left wall as in particle_x; beyond the
   * right edge xend, which moves with the upstream flow during the step,
   * x mirrors about xend and u_x about u0 (the flow's frame) */
  const int nxs = XS(p), nxe = XE(p);
  const double delx = p->delx, c = p->c;
  const double xend = nxe * delx + u0 / sqrt(1 + (u0 * u0) / (c * c)) * p->delt;
  for (int isp = 1; isp <= p->nsp; ++isp)
    for (int j = p->nys; j <= p->nye; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        double *u = up + iup(p, 1, ii, j, isp);
        const int ipos = (int)(u[0] / delx);
        if (ipos < nxs + 1) {
          u[0] = 2. * (nxs + 1) * delx - u[0];
          u[2] = -u[2];
          u[3] = -u[3];
          u[4] = -u[4];
        } else if (u[0] > xend) {
          u[0] = 2. * xend - u[0];
          u[2] = 2. * u0 - u[2];
          u[3] = -u[3];
          u[4] = -u[4];
        }
      }
  return OR_OK;
}

int or_bc_particle_x(const or_params *p, double *up, const int *np2) {
  if (p->bc != OR_BC_PERIODIC) {
    wall_x(p, up, np2);
    return OR_OK;
  }
  /* boundary_periodic.f90:61-96, rounding toward -inf (:74). */
  const int nx = NXg(p);
  const double span = (double)nx * p->delx;
  int bad = 0;
  for (int isp = 1; isp <= p->nsp; ++isp)
#pragma omp parallel for reduction(| : bad)
    for (int j = p->nys; j <= p->nye; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        double *x = up + iup(p, 1, ii, j, isp);
        const int ipos = (int)(*x / p->delx);
        if (ipos < p->nxgs)
          *x = or_add_rd(*x, span);
        else if (ipos >= p->nxge + 1)
          *x = or_add_rd(*x, -span);
        const int inew = (int)(*x / p->delx);
        if (inew < p->nxgs || inew > p->nxge) bad = 1;
      }
  return bad ? OR_ERR_RANGE : OR_OK;
}

int or_bc_particle_y(const or_params *p, or_state *s, double *up, int *np2) {
  /* boundary_periodic.f90:99-248, single rank (nup = ndown = self),
   * one thread, rounding toward -inf (:124). */
  const int nys = p->nys, nye = p->nye, ndim = p->ndim, np = p->np;
  const int ny = NYl(p);
  const double span = (double)(p->nyge - p->nygs + 1) * p->delx;
  int *cnt = (int *)calloc((size_t)ny + 2, sizeof(int));   /* cnt(nys-1:nye+1) */
  int *cnt2 = (int *)calloc((size_t)ny, sizeof(int));      /* cnt2(nys:nye) */
#define CNT(j) cnt[(j) - (nys - 1)]
#define CNT2(j) cnt2[(j) - nys]
#define FLAG(ii, j) s->flag[(size_t)((ii) - 1) + (size_t)np * (size_t)((j) - nys)]
#define BFF(k, j) s->bff[(size_t)((k) - 1) + (size_t)ndim * (size_t)np * (size_t)((j) - (nys - 1))]
  int status = OR_OK;
  for (int isp = 1; isp <= p->nsp && status == OR_OK; ++isp) {
    for (int j = nys - 1; j <= nye + 1; ++j) CNT(j) = 0;
    for (int j = nys; j <= nye; ++j) CNT2(j) = 0;
    /* pass 1 :143-169 */
    for (int j = nys; j <= nye && status == OR_OK; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        double *rec = up + iup(p, 1, ii, j, isp);
        const int jpos = (int)(rec[1] / p->delx);
        if (jpos != j) {
          if (jpos < nys - 1 || jpos > nye + 1) { status = OR_ERR_RANGE; break; }
          if (jpos <= p->nygs - 1)
            rec[1] = or_add_rd(rec[1], span);
          else if (jpos >= p->nyge + 1)
            rec[1] = or_add_rd(rec[1], -span);
          if (CNT(jpos) >= np) { status = OR_ERR_NP_OVER; break; }
          for (int k = 1; k <= ndim; ++k) BFF(k + ndim * CNT(jpos), jpos) = rec[k - 1];
          CNT(jpos) = CNT(jpos) + 1;
          CNT2(j) = CNT2(j) + 1;
          FLAG(CNT2(j), j) = ii;
        }
      }
    if (status) break;
    /* transfer to rank-1 (self) :174-180: bff(nys-1) appended to bff(nye) */
    {
      const int ct = CNT(nys - 1);
      if (CNT(nye) + ct > np) { status = OR_ERR_NP_OVER; break; }
      memmove(&BFF(ndim * CNT(nye) + 1, nye), &BFF(1, nys - 1),
              sizeof(double) * (size_t)ndim * (size_t)ct);
      CNT(nye) = CNT(nye) + ct;
    }
    /* transfer to rank+1 (self) :183-189: bff(nye+1) appended to bff(nys) */
    {
      const int ct = CNT(nye + 1);
      if (CNT(nys) + ct > np) { status = OR_ERR_NP_OVER; break; }
      memmove(&BFF(ndim * CNT(nys) + 1, nys), &BFF(1, nye + 1),
              sizeof(double) * (size_t)ndim * (size_t)ct);
      CNT(nys) = CNT(nys) + ct;
    }
    /* pass 2 :193-226 */
    for (int j = nys; j <= nye; ++j) {
      int *n2 = &np2[inp2(p, j, isp)];
      int iii = 0;
      int cnt_tmp = CNT2(j);
      for (int ii = 1; ii <= CNT2(j); ++ii) {
        if (CNT(j) == 0) {
          if (*n2 < FLAG(ii, j)) break;
          int leave = 0;
          while (*n2 == FLAG(cnt_tmp, j)) {
            *n2 = *n2 - 1;
            if (*n2 < FLAG(ii, j)) { leave = 1; break; }
            cnt_tmp = cnt_tmp - 1;
          }
          if (leave) break;
          for (int k = 1; k <= ndim; ++k)
            up[iup(p, k, FLAG(ii, j), j, isp)] = up[iup(p, k, *n2, j, isp)];
          *n2 = *n2 - 1;
        } else {
          for (int k = 1; k <= ndim; ++k)
            up[iup(p, k, FLAG(ii, j), j, isp)] = BFF(k + ndim * iii, j);
          iii = iii + 1;
          CNT(j) = CNT(j) - 1;
        }
      }
      if (CNT(j) > 0) {
        if (*n2 + CNT(j) > np) { status = OR_ERR_NP_OVER; break; }
        for (int ii = 1; ii <= CNT(j); ++ii)
          for (int k = 1; k <= ndim; ++k)
            up[iup(p, k, *n2 + ii, j, isp)] = BFF(ndim * iii + k + ndim * (ii - 1), j);
      }
    }
    if (status) break;
    /* :228-236 */
    for (int j = nys; j <= nye; ++j) {
      np2[inp2(p, j, isp)] = np2[inp2(p, j, isp)] + CNT(j);
      if (np2[inp2(p, j, isp)] > np) status = OR_ERR_NP_OVER;
    }
  }
#undef CNT
#undef CNT2
#undef FLAG
#undef BFF
  free(cnt);
  free(cnt2);
  return status;
}

/* ---------------------------------------------------------- sort__bucket */
int or_sort_bucket(const or_params *p, double *gp, const double *up,
                   int *cumcnt, const int *np2) {
  const int nxs = XS(p), nxe = XE(p), ndim = p->ndim;
  int bad = 0;
  for (int isp = 1; isp <= p->nsp; ++isp)
#pragma omp parallel for reduction(| : bad)
    for (int j = p->nys; j <= p->nye; ++j) {
      const int n = nxe - nxs + 1;
      int *cnt = (int *)calloc((size_t)n, sizeof(int));         /* cnt(nxs:nxe) */
      int *sum_cnt = (int *)calloc((size_t)n + 1, sizeof(int)); /* sum_cnt(nxs:nxe+1) */
      const int m = np2[inp2(p, j, isp)];
      int ok = 1;
      /* sort.f90:59-62 */
      for (int ii = 1; ii <= m; ++ii) {
        const int i = (int)(up[iup(p, 1, ii, j, isp)]);
        if (i < nxs || i > nxe) { ok = 0; break; }
        cnt[i - nxs] = cnt[i - nxs] + 1;
      }
      if (ok) {
        /* :64-69 */
        sum_cnt[0] = 0;
        cumcnt[icc(p, nxs, j, isp)] = 0;
        for (int i = nxs + 1; i <= nxe + 1; ++i) {
          sum_cnt[i - nxs] = sum_cnt[i - 1 - nxs] + cnt[i - 1 - nxs];
          cumcnt[icc(p, i, j, isp)] = sum_cnt[i - nxs];
        }
        /* stable scatter :71-75 */
        for (int ii = 1; ii <= m; ++ii) {
          const int i = (int)(up[iup(p, 1, ii, j, isp)]);
          memcpy(&gp[iup(p, 1, sum_cnt[i - nxs] + 1, j, isp)],
                 &up[iup(p, 1, ii, j, isp)], sizeof(double) * (size_t)ndim);
          sum_cnt[i - nxs] = sum_cnt[i - nxs] + 1;
        }
      } else {
        bad = 1;
      }
      free(cnt);
      free(sum_cnt);
    }
  return bad ? OR_ERR_RANGE : OR_OK;
}

/* --------------------------------------------------------- energy sums */
/* --------------------------------------------------- moment diagnostics */
static inline size_t imom(const or_params *p, int k, int i, int j, int isp) {
  /* mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) */
  return (size_t)(k - 1) +
         7u * ((size_t)(i - (p->nxgs - 1)) +
               (size_t)(NXg(p) + 2) * ((size_t)(j - (p->nys - 1)) +
                                       (size_t)(NYl(p) + 2) * (size_t)(isp - 1)));
}

int or_mom_calc_accl(const or_params *p, double *gp, const double *up,
                     const double *uf, const int *cumcnt) {
  /* mom_calc.f90:48-164: particle.f90's velocity update with the step that
   * mom_calc__init keeps, delt*0.5 (mom_calc.f90:34), and no move: the
   * positions are copied.  gp(6) is not written. */
  const int nxs = XS(p), nxe = XE(p), nys = p->nys, nye = p->nye;
  const int nxt = nxe - nxs + 3, nyt = nye - nys + 3;
  const double d_delx = 1. / p->delx;           /* mom_calc.f90:41 */
  const double delt = p->delt * 0.5, c = p->c;  /* :34 */
  double *tmp = (double *)malloc(sizeof(double) * 6u * (size_t)nxt * (size_t)nyt);
  if (!tmp) return OR_ERR_ARG;
#define TMP(k, i, j) tmp[(size_t)((k) - 1) + 6u * ((size_t)((i) - (nxs - 1)) + (size_t)nxt * (size_t)((j) - (nys - 1)))]
#define UF(k, i, j) uf[iuf(p, k, i, j)]
  /* fields at (i+1/2, j+1/2): :68-81 */
  for (int j = nys - 1; j <= nye + 1; ++j)
    for (int i = nxs - 1; i <= nxe + 1; ++i) {
      TMP(1, i, j) = 0.5 * (UF(1, i, j) + UF(1, i, j + 1));
      TMP(2, i, j) = 0.5 * (UF(2, i, j) + UF(2, i + 1, j));
      TMP(3, i, j) = 0.25 * (UF(3, i, j) + UF(3, i + 1, j) + UF(3, i, j + 1) +
                             UF(3, i + 1, j + 1));
      TMP(4, i, j) = 0.5 * (UF(4, i, j) + UF(4, i + 1, j));
      TMP(5, i, j) = 0.5 * (UF(5, i, j) + UF(5, i, j + 1));
      TMP(6, i, j) = UF(6, i, j);
    }
  int bad = 0;
  for (int j = nys; j <= nye; ++j)
    for (int i = nxs; i <= nxe; ++i)
      for (int isp = 1; isp <= p->nsp; ++isp) {
        const double q = p->q[isp - 1], r = p->r[isp - 1];
        const double fac1 = q / r * 0.5 * delt;  /* :90 */
        const double txxx = fac1 * fac1;         /* :91 */
        const double fac2 = q * delt / r;        /* :92 */
        const int i0 = cumcnt[icc(p, i, j, isp)], i1 = cumcnt[icc(p, i + 1, j, isp)];
        if (i0 < 0 || i1 > p->np) { bad = 1; continue; }   /* i1 < i0: an empty range */
        for (int ii = i0 + 1; ii <= i1; ++ii) {
          const double *u = up + iup(p, 1, ii, j, isp);
          double *g = gp + iup(p, 1, ii, j, isp);
          double sh[3][2], dh;
          /* :97-105 */
          dh = u[0] * d_delx - 0.5 - i;
          sh[0][0] = 0.5 * (0.5 - dh) * (0.5 - dh);
          sh[1][0] = 0.75 - dh * dh;
          sh[2][0] = 0.5 * (0.5 + dh) * (0.5 + dh);
          dh = u[1] * d_delx - 0.5 - j;
          sh[0][1] = 0.5 * (0.5 - dh) * (0.5 - dh);
          sh[1][1] = 0.75 - dh * dh;
          sh[2][1] = 0.5 * (0.5 + dh) * (0.5 + dh);
          /* :107-129 */
          double f[6];
          for (int k = 1; k <= 6; ++k)
            f[k - 1] =
                (TMP(k, i - 1, j - 1) * sh[0][0] + TMP(k, i, j - 1) * sh[1][0] +
                 TMP(k, i + 1, j - 1) * sh[2][0]) * sh[0][1] +
                (TMP(k, i - 1, j) * sh[0][0] + TMP(k, i, j) * sh[1][0] +
                 TMP(k, i + 1, j) * sh[2][0]) * sh[1][1] +
                (TMP(k, i - 1, j + 1) * sh[0][0] + TMP(k, i, j + 1) * sh[1][0] +
                 TMP(k, i + 1, j + 1) * sh[2][0]) * sh[2][1];
          const double bpx = f[0], bpy = f[1], bpz = f[2];
          const double epx = f[3], epy = f[4], epz = f[5];
          /* :132-148 */
          double uvm1 = u[2] + fac1 * epx;
          double uvm2 = u[3] + fac1 * epy;
          double uvm3 = u[4] + fac1 * epz;
          const double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
          const double igam = 1. / gam;
          const double fac1r = fac1 * igam;
          const double fac2r =
              fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
          const double uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
          const double uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
          const double uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
          uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
          uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
          uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
          /* :151-155 */
          g[0] = u[0];
          g[1] = u[1];
          g[2] = uvm1 + fac1 * epx;
          g[3] = uvm2 + fac1 * epy;
          g[4] = uvm3 + fac1 * epz;
        }
      }
  free(tmp);
#undef TMP
#undef UF
  return bad ? OR_ERR_RANGE : OR_OK;
}

int or_mom_calc_nvt(const or_params *p, double *mom, const double *up,
                    const int *np2) {
  /* mom_calc.f90:167-249, in its serial order (species, row, particle).
   * ih = int(x/delx - 0.5) truncates toward zero, as Fortran's int() does:
   * below x = delx/2 of cell 0 it is 0, not -1, and the weights extrapolate.
   * A particle whose nodes fall outside mom's bounds is skipped and reported
   * (the reference would write out of bounds). */
  const double d_delx = 1. / p->delx, c = p->c;
  const size_t n = 7u * (size_t)(NXg(p) + 2) * (size_t)(NYl(p) + 2) * (size_t)p->nsp;
  for (size_t t = 0; t < n; ++t) mom[t] = 0.0;   /* :182 */
  int bad = 0;
  for (int isp = 1; isp <= p->nsp; ++isp)
    for (int j = p->nys; j <= p->nye; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        const double *u = up + iup(p, 1, ii, j, isp);
        const int ih = (int)(u[0] * d_delx - 0.5);   /* :190-191 */
        const int jh = (int)(u[1] * d_delx - 0.5);
        if (ih < p->nxgs - 1 || ih > p->nxge || jh < p->nys - 1 || jh > p->nye) {
          bad = 1;
          continue;
        }
        const double dx = u[0] * d_delx - 0.5 - ih;  /* :193-196 */
        const double dxm = 1. - dx;
        const double dy = u[1] * d_delx - 0.5 - jh;
        const double dym = 1. - dy;
        const double gam =
            1. / sqrt(1.0 + (u[2] * u[2] + u[3] * u[3] + u[4] * u[4]) / (c * c));
        /* per node (ih, jh), (ih+1, jh), (ih, jh+1), (ih+1, jh+1): :203-243 */
        const double wx[2] = {dxm, dx}, wy[2] = {dym, dy};
        const double v[7] = {1.0, u[2] * gam, u[3] * gam, u[4] * gam,
                             u[2] * u[2] * gam, u[3] * u[3] * gam, u[4] * u[4] * gam};
        for (int k = 1; k <= 7; ++k)
          for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
              /* N: dx*dy; the others: ((u gam) dx) dy, as written */
              const double w = k == 1 ? wx[a] * wy[b] : v[k - 1] * wx[a] * wy[b];
              mom[imom(p, k, ih + a, jh + b, isp)] += w;
            }
      }
  return bad ? OR_ERR_RANGE : OR_OK;
}

void or_bc_mom(const or_params *p, double *mom) {
  /* boundary_periodic.f90:571-636 with one rank (nup = ndown = itself): the
   * x ghosts fold into the opposite edge over every row, ghost rows
   * included (:579-584); then, per species, row nys-1 into row nye (:586-608)
   * and row nye+1 into row nys (:611-633), ghost columns included. */
  const int nxgs = p->nxgs, nxge = p->nxge, nys = p->nys, nye = p->nye;
  const int wall = p->bc != OR_BC_PERIODIC;
  /* the wall families fold each x ghost into its own edge
   * (boundary_shock.f90:636-641, boundary_reconnection.f90:590-595) */
  for (int isp = 1; isp <= p->nsp; ++isp)
    for (int j = nys - 1; j <= nye + 1; ++j)
      for (int k = 1; k <= 7; ++k) {
        mom[imom(p, k, nxgs, j, isp)] += mom[imom(p, k, wall ? nxgs - 1 : nxge + 1, j, isp)];
        mom[imom(p, k, nxge, j, isp)] += mom[imom(p, k, wall ? nxge + 1 : nxgs - 1, j, isp)];
      }
  for (int isp = 1; isp <= p->nsp; ++isp) {
    for (int i = nxgs - 1; i <= nxge + 1; ++i)
      for (int k = 1; k <= 7; ++k)
        mom[imom(p, k, i, nye, isp)] += mom[imom(p, k, i, nys - 1, isp)];
    for (int i = nxgs - 1; i <= nxge + 1; ++i)
      for (int k = 1; k <= 7; ++k)
        mom[imom(p, k, i, nys, isp)] += mom[imom(p, k, i, nye + 1, isp)];
  }
}

void or_energy(const or_params *p, const double *up, const double *uf,
               const int *np2, double *out) {
  /* proj/weibel/app.f90:499-535 */
  const double c = p->c;
  for (int isp = 1; isp <= p->nsp; ++isp) {
    double vene = 0.0;
    for (int j = p->nys; j <= p->nye; ++j)
      for (int ii = 1; ii <= np2[inp2(p, j, isp)]; ++ii) {
        const double *u = up + iup(p, 1, ii, j, isp);
        const double u2 = u[2] * u[2] + u[3] * u[3] + u[4] * u[4];
        const double gam = sqrt(1 + u2 / (c * c));
        vene = vene + p->r[isp - 1] * (gam - 1);
      }
    out[isp - 1] = vene;
  }
  double bfield = 0.0, efield = 0.0;
  for (int j = p->nys; j <= p->nye; ++j)
    for (int i = p->nxgs; i <= p->nxge; ++i) {
      const double *f = uf + iuf(p, 1, i, j);
      bfield = bfield + f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
      efield = efield + f[3] * f[3] + f[4] * f[4] + f[5] * f[5];
    }
  out[p->nsp] = efield / (8 * OR_PI);
  out[p->nsp + 1] = bfield / (8 * OR_PI);
}

/* ----------------------------------------------------------------- step */
int or_step(const or_params *p, or_state *s, double *up, double *gp,
            double *uf, int *cumcnt, int *np2) {
  /* proj/weibel/app.f90:102-106, proj/reconnection/app.f90:101-106; the
   * shock's injection edge comes between the push and the current
   * (proj/shock/app.f90:112-117) and it has no particle_x */
  int st = or_particle_solv(p, gp, up, uf, cumcnt);
  if (st) return st;
  if (p->bc == OR_BC_SHOCK) or_bc_injection(p, gp, np2, p->u0);
  st = or_field_fdtd_i(p, s, uf, up, gp, cumcnt);
  if (st) return st;
  if (p->bc != OR_BC_SHOCK) st = or_bc_particle_x(p, gp, np2);
  if (st) return st;
  st = or_bc_particle_y(p, s, gp, np2);
  if (st) return st;
  return or_sort_bucket(p, up, gp, cumcnt, np2);
}

/* Threads used by the OpenMP loops (parity runs set 1). */
void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
