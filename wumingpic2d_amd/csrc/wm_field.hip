// wm_field.hip -- implicit-FDTD field solve (field.f90:66-186, 319-461) and
// the field/current/potential boundaries: periodic (boundary_periodic.f90)
// and the conducting walls of the shock and reconnection families
// (proj/shock/boundary_shock.f90, proj/reconnection/boundary_reconnection.f90).
//
// Fields are component planes of (nx+4) x (ny+4), x fastest.  The CG works
// directly on the interior of the delta-B planes of df (which also carry the
// previous step's solution, the reference's warm start field.f90:98,355) and
// takes the 1-cell halo of set_boundary_phi by index: the periodic wrap, or
// the wall ghosts.  The wall families solve over the active columns
// [nxs, nxe] only (the shock's box grows, proj/shock/app.f90:613-690).
// The three Helmholtz solves are batched in one launch per sweep
// (blockIdx.y = component); components that have converged skip their work.
// Dot products are reduced deterministically: fixed per-workgroup trees, then
// the last workgroup sums the partials in a fixed order and performs the
// scalar update of the reference on the device.
#include "wm_internal.h"

#include <math.h>

namespace {

constexpr double PI = 3.141592653589793115997963468544185161590576171875;  // field.f90:13

struct FC {      // field constants, field.f90:53-57
  double f1, f2, f3, f4, f5, gfac, delt;
};

FC make_fc(const wm_grid &g) {
  FC f;
  f.f1 = g.c * g.delt / g.delx;
  f.f2 = g.gfac * f.f1 * f.f1;
  f.f3 = 4.0 * PI * g.delx / g.c;
  const double t = g.delx / (g.c * g.delt * g.gfac);
  f.f4 = 4.0 + t * t;
  f.f5 = t * t;
  f.gfac = g.gfac;
  f.delt = g.delt;
  return f;
}

// ------------------------------------------------------------ boundaries
// boundary_periodic__curre (boundary_periodic.f90:357-506), row phase: one
// thread per (column, component) runs the fold/copy sequence of that column,
// for the ghost-inclusive columns c0 .. c0 + cnt - 1 (nxs-2 .. nxe+2).
__global__ void k_curre_rows(FieldGeom fg, double *uj, int c0, int cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt * 3) return;
  const int li = c0 + t % cnt, comp = t / cnt;
  double *U = uj + (size_t)comp * fg.ps + li;
  const size_t gx = fg.gx;
  const int ny = fg.ny;
  // rows: nys-2 -> 0, nys-1 -> 1, nys -> 2, nys+1 -> 3,
  //       nye-1 -> ny, nye -> ny+1, nye+1 -> ny+2, nye+2 -> ny+3
  const double r1 = U[0], r2 = U[gx];
  U[ny * gx] = U[ny * gx] + r1;                 // :391-396
  U[(ny + 1) * gx] = U[(ny + 1) * gx] + r2;
  const double s1 = U[(ny + 2) * gx], s2 = U[(ny + 3) * gx];
  U[2 * gx] = U[2 * gx] + s1;                   // :422-427
  U[3 * gx] = U[3 * gx] + s2;
  const double t1 = U[2 * gx], t2 = U[3 * gx];
  U[(ny + 2) * gx] = t1;                        // :455-460
  U[(ny + 3) * gx] = t2;
  const double v1 = U[ny * gx], v2 = U[(ny + 1) * gx];
  U[0] = v1;                                    // :486-491
  U[gx] = v2;
}

// Column phase (boundary_periodic.f90:495-506), one thread per (row, comp).
__global__ void k_curre_cols(FieldGeom fg, double *uj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fg.gy * 3) return;
  const int lj = t % fg.gy, comp = t / fg.gy;
  double *R = uj + (size_t)comp * fg.ps + (size_t)lj * fg.gx;
  const int nx = fg.nx;
  // cols: nxs-2 -> 0, nxs-1 -> 1, nxs -> 2, nxs+1 -> 3,
  //       nxe-1 -> nx, nxe -> nx+1, nxe+1 -> nx+2, nxe+2 -> nx+3
  R[nx] = R[nx] + R[0];
  R[nx + 1] = R[nx + 1] + R[1];
  R[2] = R[2] + R[nx + 2];
  R[3] = R[3] + R[nx + 3];
  R[0] = R[nx];
  R[1] = R[nx + 1];
  R[nx + 2] = R[2];
  R[nx + 3] = R[3];
}

// boundary_periodic__dfield (boundary_periodic.f90:263-345): ghost rows of
// the interior columns x0 .. x0 + cnt - 1 (all of them, or the wall
// families' active ones, boundary_shock.f90:326-392).
__global__ void k_dfield_rows(FieldGeom fg, double *df, int ncomp, int x0, int cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt * ncomp) return;
  const int li = 2 + x0 + t % cnt, comp = t / cnt;
  double *U = df + (size_t)comp * fg.ps + li;
  const size_t gx = fg.gx;
  const int ny = fg.ny;
  U[(ny + 2) * gx] = U[2 * gx];
  U[(ny + 3) * gx] = U[3 * gx];
  U[0] = U[ny * gx];
  U[gx] = U[(ny + 1) * gx];
}

// boundary_periodic.f90:347-352: ghost columns of every row.
__global__ void k_dfield_cols(FieldGeom fg, double *df, int ncomp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fg.gy * ncomp) return;
  const int lj = t % fg.gy, comp = t / fg.gy;
  double *R = df + (size_t)comp * fg.ps + (size_t)lj * fg.gx;
  const int nx = fg.nx;
  R[0] = R[nx];
  R[1] = R[nx + 1];
  R[nx + 2] = R[2];
  R[nx + 3] = R[3];
}

// The conducting walls' ghost columns of df, every row (after the rows):
// the shock's (boundary_shock.f90:396-405) -- Bx, Ey, Ez odd about the wall
// at nxs - 1/2, By, Bz, Ex even about column nxs, zero at nxe+1 -- or the
// reconnection's (boundary_reconnection.f90:350-359), whose right wall is
// the mirror image at nxe - 1/2 and rewrites column nxe of Bx, Ey, Ez.
__global__ void k_dfield_walls(FieldGeom fg, double *df, XSide X) {
  const int lj = blockIdx.x * blockDim.x + threadIdx.x;
  if (lj >= fg.gy) return;
  double *R = df + (size_t)lj * fg.gx;
  const size_t ps = fg.ps;
  const int L = X.xa + 2, E = X.xb + 2;      // ghost-inclusive columns of nxs, nxe
  R[L - 1] = -R[L];
  for (int k = 1; k <= 3; ++k) R[k * ps + L - 1] = R[k * ps + L + 1];
  for (int k = 4; k <= 5; ++k) R[k * ps + L - 1] = -R[k * ps + L];
  if (X.wall == WM_BC_SHOCK) {
    for (int k = 0; k < 6; ++k) R[k * ps + E + 1] = 0.0;
  } else {
    R[E] = -R[E - 1];
    for (int k = 1; k <= 3; ++k) R[k * ps + E + 1] = R[k * ps + E - 1];
    for (int k = 4; k <= 5; ++k) R[k * ps + E] = -R[k * ps + E - 1];
  }
}

// boundary_periodic__phi (boundary_periodic.f90:511-568) on an
// (nx+2) x (ny+2) array: rows, then columns.
__global__ void k_phi_rows(int nx, int ny, double *phi) {
  const int a = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (a > nx) return;
  const size_t w = nx + 2;
  phi[a + (size_t)(ny + 1) * w] = phi[a + w];
  phi[a] = phi[a + (size_t)ny * w];
}
__global__ void k_phi_cols(int nx, int ny, double *phi) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > ny + 1) return;
  const size_t w = nx + 2;
  phi[(size_t)b * w] = phi[(size_t)b * w + nx];
  phi[(size_t)b * w + nx + 1] = phi[(size_t)b * w + 1];
}
// The wall families' columns of phi(nxs-1:nxe+1, ...) for component l
// (boundary_shock.f90:603-623, boundary_reconnection.f90:557-577).
__global__ void k_phi_walls(int nx, int ny, double *phi, int l) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > ny + 1) return;
  double *R = phi + (size_t)b * (nx + 2);
  R[0] = l == 1 ? -R[1] : R[2];
  R[nx + 1] = 0.0;
}

// ----------------------------------------------------------- RHS, dE, +=
// gkl (field.f90:125-146), interior cells, gkl planes nx x ny.
__global__ void k_rhs(FieldGeom fg, FC f, XSide X, const double *__restrict__ uf,
                      const double *__restrict__ uj, double *__restrict__ gkl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fg.nx * fg.ny) return;
  const int ii = t % fg.nx, jj = t / fg.nx;
  if (ii < X.xa || ii > X.xb) return;
  const size_t gx = fg.gx, ps = fg.ps;
  const size_t c = (size_t)(ii + 2) + gx * (size_t)(jj + 2);
  const double *B1 = uf, *B2 = uf + ps, *B3 = uf + 2 * ps;
  const double *E1 = uf + 3 * ps, *E2 = uf + 4 * ps, *E3 = uf + 5 * ps;
  const double *J1 = uj, *J2 = uj + ps, *J3 = uj + 2 * ps;
  const size_t n = (size_t)fg.nx * fg.ny;
  gkl[t] = f.f2 * (B1[c - gx] + B1[c - 1] - 4. * B1[c] + B1[c + 1] + B1[c + gx] +
                   f.f3 * (-J3[c - gx] + J3[c])) -
           f.f1 * (-E3[c - gx] + E3[c]);
  gkl[n + t] = f.f2 * (B2[c - gx] + B2[c - 1] - 4. * B2[c] + B2[c + 1] + B2[c + gx] -
                       f.f3 * (-J3[c - 1] + J3[c])) +
               f.f1 * (-E3[c - 1] + E3[c]);
  gkl[2 * n + t] =
      f.f2 * (B3[c - gx] + B3[c - 1] - 4. * B3[c] + B3[c + 1] + B3[c + gx] +
              f.f3 * (-J2[c - 1] + J2[c] + J1[c - gx] - J1[c])) -
      f.f1 * (-E2[c - 1] + E2[c] + E1[c - gx] - E1[c]);
}

// delta-E (field.f90:154-171), interior cells.
__global__ void k_delta_e(FieldGeom fg, FC f, XSide X, const double *__restrict__ uf,
                          const double *__restrict__ uj, double *__restrict__ df) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fg.nx * fg.ny) return;
  const int ii = t % fg.nx, jj = t / fg.nx;
  if (ii < X.xa || ii > X.xb) return;
  const size_t gx = fg.gx, ps = fg.ps;
  const size_t c = (size_t)(ii + 2) + gx * (size_t)(jj + 2);
  const double *dB1 = df, *dB2 = df + ps, *dB3 = df + 2 * ps;
  const double *B1 = uf, *B2 = uf + ps, *B3 = uf + 2 * ps;
  const double *J1 = uj, *J2 = uj + ps, *J3 = uj + 2 * ps;
  const double fpt = 4. * PI * f.delt;
  df[3 * ps + c] = f.f1 * (f.gfac * (-dB3[c] + dB3[c + gx]) + (-B3[c] + B3[c + gx])) - fpt * J1[c];
  df[4 * ps + c] = -f.f1 * (f.gfac * (-dB3[c] + dB3[c + 1]) + (-B3[c] + B3[c + 1])) - fpt * J2[c];
  df[5 * ps + c] = f.f1 * (f.gfac * (-dB2[c] + dB2[c + 1] + dB1[c] - dB1[c + gx]) +
                           (-B2[c] + B2[c + 1] + B1[c] - B1[c + gx])) -
                   fpt * J3[c];
}

// uf += df over every cell including ghosts (field.f90:176-184).
__global__ void k_add(size_t n, double *__restrict__ uf, const double *__restrict__ df) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) uf[t] = uf[t] + df[t];
}
// The same over columns nxs-2 .. nxe+2 only (an active range short of the
// grid): ghost-inclusive columns xa .. xb + 4 of every row and component.
__global__ void k_add_cols(FieldGeom fg, XSide X, double *__restrict__ uf,
                           const double *__restrict__ df) {
  const int w = X.xb - X.xa + 5;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)6 * fg.gy * w) return;
  const int a = (int)(t % w);
  const size_t rc = t / w;                     // component * gy + row
  const size_t q = (rc / fg.gy) * fg.ps + (rc % fg.gy) * fg.gx + X.xa + a;
  uf[q] = uf[q] + df[q];
}

// -------------------------------------------------------------------- CG
constexpr int CGB = 256;   // threads per CG workgroup

struct CgArgs {
  int nx, ny, gx, nblk;    // nblk = workgroups per component
  int nbx;                 // workgroups per row
  size_t ps;               // df plane stride
  size_t n;                // nx*ny
  double f4, f5, err;
  int ite_max;
  XSide X;                 // periodic, or the walls and the active columns
};

// Fixed-order workgroup sum of NV values: a butterfly within each wave, then
// the CGB / 64 wave totals in order; the totals land in every thread.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double (*sh)[CGB]) {
  constexpr int NW = CGB / 64;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double w = wm_wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) sh[k][wave] = w;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = sh[k][0];
#pragma unroll
    for (int q = 1; q < NW; ++q) t += sh[k][q];
    v[k] = t;
  }
}

// Publish this workgroup's NV partial sums (fixed-order tree).
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double (*sh)[CGB], double *partials,
                                      int nblk, int bid) {
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * nblk + bid] = v[k];
  }
}

// Fixed-order total of NV partial arrays by one workgroup.
template <int NV>
__device__ inline void partial_total(const double *partials, int nblk, double (*sh)[CGB],
                                     double (&tot)[NV]) {
  double a[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) a[k] = 0.0;
  // unrolled so that the loads go out together; the additions keep their order
#pragma unroll 8
  for (int b = threadIdx.x; b < nblk; b += CGB) {
#pragma unroll
    for (int k = 0; k < NV; ++k) a[k] += partials[(size_t)k * nblk + b];
  }
  block_sum<NV>(a, sh);
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = a[k];
}

__device__ inline int wrapm(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// XCD-aware workgroup order: the dispatcher deals workgroups round-robin to
// the 8 XCDs; renumber them so that each XCD (own L2) sweeps a contiguous band
// of rows and the stencil's neighbour rows hit the same L2.
__device__ inline int xcd_order(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = bid & 7, s = bid >> 3;
  return x * q + min(x, r) + s;
}

// Interior cell of this thread in the row-blocked CG launch: workgroup b
// covers x in [(b % nbx)*CGB, +CGB) of row b / nbx.
struct CgCell {
  int ii, jj, t, b;
  bool in;
};
__device__ inline CgCell cg_cell(const CgArgs &A) {
  CgCell c;
  c.b = xcd_order(blockIdx.x, A.nblk);
  c.jj = c.b / A.nbx;
  c.ii = (c.b - c.jj * A.nbx) * CGB + threadIdx.x;
  c.in = c.ii >= A.X.xa && c.ii <= A.X.xb;
  c.t = c.jj * A.nx + c.ii;
  return c;
}

// Value at column a, row b of component l of a CG field v(x, y): the
// periodic wrap, or the wall ghosts of set_boundary_phi
// (boundary_shock.f90:603-623): left of the active columns -v(xa) for Bx
// and v(xa + 1) for By and Bz, right of them zero.
template <class F>
__device__ inline double cg_at(const CgArgs &A, int l, int a, int b, F v) {
  const int bw = wrapm(b, A.ny);
  if (!A.X.wall) return v(wrapm(a, A.nx), bw);
  if (a < A.X.xa) return l == 0 ? -v(A.X.xa, bw) : v(A.X.xa + 1, bw);
  if (a > A.X.xb) return 0.0;
  return v(a, bw);
}

// Initial residual (field.f90:350-385): b = f5*gkl, r = b + N4(phi) -
// f4*phi with the periodic halo of boundary_periodic__phi, p = r; partial
// sums of b.b and r.r.
__global__ __launch_bounds__(CGB) void k_cg_init(CgArgs A, const double *__restrict__ gkl,
                                                 const double *__restrict__ df,
                                                 double *__restrict__ r, double *__restrict__ p,
                                                 double *partials,
                                                 double *__restrict__ x0 = nullptr) {
  __shared__ double sh[2][CGB];
  const int l = blockIdx.y;
  const CgCell C = cg_cell(A);
  const int ii = C.ii, jj = C.jj, t = C.t;
  double v[2] = {0.0, 0.0};
  if (C.in) {
    const double *phi = df + (size_t)l * A.ps;
    auto PH = [&](int a, int b) {
      return cg_at(A, l, a, b, [&](int x, int y) {
        return phi[(size_t)(x + 2) + (size_t)A.gx * (size_t)(y + 2)];
      });
    };
    const double b = A.f5 * gkl[(size_t)l * A.n + t];
    const double rr = b + PH(ii, jj - 1) + PH(ii - 1, jj) - A.f4 * PH(ii, jj) +
                      PH(ii + 1, jj) + PH(ii, jj + 1);
    r[(size_t)l * A.n + t] = rr;
    p[(size_t)l * A.n + t] = rr;
    if (x0) x0[(size_t)l * A.n + t] = PH(ii, jj);   // the next flexible solve's warm start
    v[0] = b * b;
    v[1] = rr * rr;
  }
  block_partials<2>(v, sh, partials + (size_t)l * 2 * A.nblk, A.nblk, C.b);
}

// A*p = -p(i,j-1) - p(i-1,j) + f4*p(i,j) - p(i+1,j) - p(i,j+1) (field.f90:400-402),
// with one fixed rounding sequence so every evaluation of a cell agrees.
__device__ inline double apply_a(double f4, double pc, double ps_, double pw, double pe,
                                 double pn) {
  double a = __dsub_rn(-ps_, pw);
  a = __fma_rn(f4, pc, a);
  a = __dsub_rn(a, pe);
  return __dsub_rn(a, pn);
}

// Direction update p = r + bv*p (field.f90:444-450) fused into the next
// iteration's p.Ap sweep (field.f90:395-415): p_new is formed on the fly at
// the five stencil points, stored, and dotted with A p_new.  bv = 0 on the
// first iteration (p = r from the initial residual).
__global__ __launch_bounds__(CGB) void k_cg_pap(CgArgs A, const double *__restrict__ r,
                                                const double *__restrict__ pold,
                                                double *__restrict__ pnew, double *partials,
                                                const CgScal *S) {
  __shared__ double sh[1][CGB];
  const int l = blockIdx.y;
  if (!S[l].active) return;
  const double bv = S[l].bv;
  const CgCell C = cg_cell(A);
  const int ii = C.ii, jj = C.jj, t = C.t;
  double v[1] = {0.0};
  if (C.in) {
    const double *R = r + (size_t)l * A.n, *P = pold + (size_t)l * A.n;
    auto PN = [&](int a, int b) {
      return cg_at(A, l, a, b, [&](int x, int y) {
        const int q = x + A.nx * y;
        return __fma_rn(bv, P[q], R[q]);
      });
    };
    const double pc = PN(ii, jj);
    const double ap = apply_a(A.f4, pc, PN(ii, jj - 1), PN(ii - 1, jj), PN(ii + 1, jj), PN(ii, jj + 1));
    pnew[(size_t)l * A.n + t] = pc;
    v[0] = pc * ap;
  }
  block_partials<1>(v, sh, partials + (size_t)l * 2 * A.nblk, A.nblk, C.b);
}

// phi += av*p, r -= av*A p, partial |r|^2 (field.f90:417-439).
__global__ __launch_bounds__(CGB) void k_cg_update(CgArgs A, double *__restrict__ df,
                                                   double *__restrict__ r,
                                                   const double *__restrict__ p,
                                                   double *partials, const CgScal *S) {
  __shared__ double sh[1][CGB];
  const int l = blockIdx.y;
  if (!S[l].active) return;
  const double av = S[l].av;
  const CgCell C = cg_cell(A);
  const int ii = C.ii, jj = C.jj, t = C.t;
  double v[1] = {0.0};
  if (C.in) {
    const double *P = p + (size_t)l * A.n;
    auto PP = [&](int a, int b) {
      return cg_at(A, l, a, b, [&](int x, int y) { return P[x + A.nx * y]; });
    };
    const double pc = P[t];
    const double ap = apply_a(A.f4, pc, PP(ii, jj - 1), PP(ii - 1, jj), PP(ii + 1, jj), PP(ii, jj + 1));
    double *phi = df + (size_t)l * A.ps + (size_t)(ii + 2) + (size_t)A.gx * (size_t)(jj + 2);
    *phi = *phi + av * pc;
    double *rp = r + (size_t)l * A.n + t;
    const double rn = *rp - av * ap;
    *rp = rn;
    v[0] = rn * rn;
  }
  block_partials<1>(v, sh, partials + (size_t)l * 2 * A.nblk, A.nblk, C.b);
}

// Scalar updates of the reference's CG, one workgroup per component.
// stage 0: after k_cg_init; sum_g = |b|^2, eps = sqrt(sum_g)*err, sumr =
//   |r|^2, run while sqrt(sumr) > eps and sum_g > eps (field.f90:356-387).
// stage 1: after k_cg_pap; ite += 1, av = sumr/pAp, sum_g = sqrt(sumr) (the
//   pre-update residual, :415, :426), error at ite_max (:427).
// stage 2: after k_cg_update; bv = |r_new|^2/sumr (:442), loop test (:387).
__global__ __launch_bounds__(CGB) void k_cg_fin(CgArgs A, int stage, const double *partials,
                                                CgScal *S) {
  __shared__ double sh[2][CGB];
  const int l = blockIdx.x;
  if (stage != 0 && !S[l].active) return;
  const double *pl = partials + (size_t)l * 2 * A.nblk;
  if (stage == 0) {
    double tot[2];
    partial_total<2>(pl, A.nblk, sh, tot);
    if (threadIdx.x == 0) {
      CgScal s;
      s.sum_g = tot[0];
      s.eps = sqrt(tot[0]) * A.err;
      s.sumr = tot[1];
      s.sum2 = 0.0;
      s.av = 0.0;
      s.bv = 0.0;
      s.ite = 0;
      s.err = 0;
      s.active = (sqrt(s.sumr) > s.eps) && (s.sum_g > s.eps);
      s.pad = 0;
      S[l] = s;
    }
  } else {
    double tot[1];
    partial_total<1>(pl, A.nblk, sh, tot);
    if (threadIdx.x == 0) {
      CgScal s = S[l];
      if (stage == 1) {
        s.ite = s.ite + 1;
        s.sum2 = tot[0];
        s.av = s.sumr / s.sum2;
        s.sum_g = sqrt(s.sumr);
        if (s.ite >= A.ite_max) s.err = 1;
      } else {
        s.bv = tot[0] / s.sumr;
        s.sumr = tot[0];
        s.active = (s.sum_g > s.eps) && !s.err;
      }
      S[l] = s;
    }
  }
}

// ------------------------------------------------------- flexible CG
// Flexible CG (Notay 2000; Polak-Ribiere beta) with the learned
// preconditioner z = M(r).  Per component:
//   z = M(r); rz = r.z; beta = -alpha_prev (z.q_prev) / rz_prev
//   [= z.(r - r_prev) / rz_prev]; p = z + beta p; q = A p; alpha = rz / p.q;
//   phi += alpha p; r -= alpha q; stop when |r| <= err |b|.
// A non-finite or non-positive curvature marks the component bad: its phi is
// restored from the warm start and the plain solve takes over.

__global__ __launch_bounds__(CGB) void k_fcg_save(CgArgs A, const double *__restrict__ df,
                                                  double *__restrict__ x0) {
  const int l = blockIdx.y;
  const CgCell C = cg_cell(A);
  if (!C.in) return;
  x0[(size_t)l * A.n + C.t] = df[(size_t)l * A.ps + (size_t)(C.ii + 2) + (size_t)A.gx * (C.jj + 2)];
}

__global__ __launch_bounds__(CGB) void k_fcg_restore(CgArgs A, double *__restrict__ df,
                                                     const double *__restrict__ x0,
                                                     const FcgScal *S) {
  const int l = blockIdx.y;
  if (!S[l].bad) return;
  const CgCell C = cg_cell(A);
  if (!C.in) return;
  df[(size_t)l * A.ps + (size_t)(C.ii + 2) + (size_t)A.gx * (C.jj + 2)] = x0[(size_t)l * A.n + C.t];
}

// p = z + beta p_old at the five stencil points, q = A p, partial p.q
__global__ __launch_bounds__(CGB) void k_fcg_dir(CgArgs A, const double *__restrict__ z,
                                                 const double *__restrict__ pold,
                                                 double *__restrict__ pnew, double *__restrict__ q,
                                                 double *partials, const FcgScal *S) {
  __shared__ double sh[1][CGB];
  const int l = blockIdx.y;
  if (!S[l].active) return;
  const bool first = S[l].it == 0;
  const double bv = S[l].beta;
  const CgCell C = cg_cell(A);
  const int ii = C.ii, jj = C.jj, t = C.t;
  double v[1] = {0.0};
  if (C.in) {
    const double *Z = z + (size_t)l * A.n, *P = pold + (size_t)l * A.n;
    auto PN = [&](int a, int b) {
      const int k = wrapm(a, A.nx) + A.nx * wrapm(b, A.ny);
      return first ? Z[k] : __fma_rn(bv, P[k], Z[k]);
    };
    const double pc = PN(ii, jj);
    const double ap = apply_a(A.f4, pc, PN(ii, jj - 1), PN(ii - 1, jj), PN(ii + 1, jj), PN(ii, jj + 1));
    pnew[(size_t)l * A.n + t] = pc;
    q[(size_t)l * A.n + t] = ap;
    v[0] = pc * ap;
  }
  block_partials<1>(v, sh, partials + (size_t)l * 2 * A.nblk, A.nblk, C.b);
}

// phi += alpha p, r -= alpha q, partial |r|^2
__global__ __launch_bounds__(CGB) void k_fcg_update(CgArgs A, double *__restrict__ df,
                                                    double *__restrict__ r,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ q,
                                                    double *partials, const FcgScal *S) {
  __shared__ double sh[1][CGB];
  const int l = blockIdx.y;
  if (!S[l].active) return;
  const double al = S[l].alpha;
  const CgCell C = cg_cell(A);
  double v[1] = {0.0};
  if (C.in) {
    const size_t o = (size_t)l * A.n + C.t;
    double *phi = df + (size_t)l * A.ps + (size_t)(C.ii + 2) + (size_t)A.gx * (size_t)(C.jj + 2);
    *phi = *phi + al * p[o];
    const double rn = r[o] - al * q[o];
    r[o] = rn;
    v[0] = rn * rn;
  }
  block_partials<1>(v, sh, partials + (size_t)l * 2 * A.nblk, A.nblk, C.b);
}

// Scalars of flexible CG; sc gets the residual scales s = |r|/sqrt(n), 1/s
// that normalise the preconditioner's input.
// stage 0 after k_cg_init, 1 after k_fcg_dots, 2 after k_fcg_dir,
// 3 after k_fcg_update.
__global__ __launch_bounds__(CGB) void k_fcg_fin(CgArgs A, int stage, const double *partials,
                                                 FcgScal *S, double *sc, int itmax) {
  __shared__ double sh[2][CGB];
  const int l = blockIdx.x;
  if (stage != 0 && !S[l].active) return;
  const double *pl = partials + (size_t)l * 2 * A.nblk;
  double tot[2] = {0.0, 0.0};
  if (stage <= 1) {
    partial_total<2>(pl, A.nblk, sh, tot);
  } else {
    double t1[1];
    partial_total<1>(pl, A.nblk, sh, t1);
    tot[0] = t1[0];
  }
  if (threadIdx.x != 0) return;
  FcgScal s = stage == 0 ? FcgScal{} : S[l];
  if (stage == 0) {
    s.eps = sqrt(tot[0]) * A.err;
    s.sumr = tot[1];
    s.it = 0;
    s.bad = 0;
    s.itmax = itmax;
    s.active = sqrt(s.sumr) > s.eps;
  } else if (stage == 1) {
    s.rz = tot[0];
    s.zq = tot[1];
    s.beta = s.it == 0 ? 0.0 : -s.alpha * s.zq / s.rz_prev;
    if (!isfinite(s.rz) || !isfinite(s.beta)) s.bad = 1;
  } else if (stage == 2) {
    s.pq = tot[0];
    s.alpha = s.rz / s.pq;
    if (!(s.pq > 0.0) || !isfinite(s.alpha)) s.bad = 1;
  } else {
    s.sumr = tot[0];
    s.rz_prev = s.rz;
    s.it += 1;
    if (!isfinite(s.sumr)) s.bad = 1;
    s.active = sqrt(s.sumr) > s.eps && s.it < s.itmax;
  }
  if (s.bad) s.active = 0;
  double sd = sqrt(s.sumr / (double)A.n);
  if (!(sd > 0.0) || !isfinite(sd)) sd = 1.0;
  sc[l] = sd;
  sc[3 + l] = 1.0 / sd;
  S[l] = s;
}

// Energy sums (proj/weibel/app.f90:499-535): per-workgroup partials.
__global__ __launch_bounds__(256) void k_energy_kin(int ny, int cap, int nsp_used,
                                                    const int *np2, const double *ux,
                                                    const double *uy, const double *uz,
                                                    double c, double *part, int nchunk) {
  __shared__ double sh[1][CGB];
  const int seg = blockIdx.y;
  const int n = np2[seg];
  double v[1] = {0.0};
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += 256 * nchunk) {
    const size_t p = (size_t)seg * cap + k;
    const double u2 = ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p];
    v[0] += sqrt(1 + u2 / (c * c)) - 1;
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[(size_t)seg * nchunk + blockIdx.x] = v[0];
  (void)ny;
  (void)nsp_used;
}

__global__ __launch_bounds__(256) void k_energy_fld(FieldGeom fg, const double *uf,
                                                    double *part) {
  __shared__ double sh[2][CGB];
  const int jj = blockIdx.x;   // one workgroup per interior row
  double v[2] = {0.0, 0.0};
  for (int ii = threadIdx.x; ii < fg.nx; ii += 256) {
    const size_t c = (size_t)(ii + 2) + (size_t)fg.gx * (size_t)(jj + 2);
    const double b1 = uf[c], b2 = uf[fg.ps + c], b3 = uf[2 * fg.ps + c];
    const double e1 = uf[3 * fg.ps + c], e2 = uf[4 * fg.ps + c], e3 = uf[5 * fg.ps + c];
    v[0] += b1 * b1 + b2 * b2 + b3 * b3;
    v[1] += e1 * e1 + e2 * e2 + e3 * e3;
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    part[2 * jj] = v[0];
    part[2 * jj + 1] = v[1];
  }
}

}  // namespace

// -------------------------------------------------------------- launchers
XSide wm_xside(const wm_ctx *ctx) {
  XSide X;
  X.wall = ctx->bcf;
  X.xa = ctx->xe > 0 ? ctx->xs - ctx->g.nxgs : 0;
  X.xb = ctx->xe > 0 ? ctx->xe - ctx->g.nxgs : ctx->g.nx - 1;
  return X;
}

// The wall families fold the current in y only, over columns nxs-2 ..
// nxe+2 (boundary_shock.f90:410-548).
int wm_k_bc_curre(wm_ctx *ctx, double *uj) {
  const FieldGeom &fg = ctx->fg;
  const XSide X = wm_xside(ctx);
  const int cnt = X.xb - X.xa + 5;
  hipLaunchKernelGGL(k_curre_rows, dim3(wm_cdiv(cnt * 3, 256)), dim3(256), 0, ctx->stream, fg, uj,
                     X.xa, cnt);
  if (ctx->bcf == WM_BC_PERIODIC)
    hipLaunchKernelGGL(k_curre_cols, dim3(wm_cdiv(fg.gy * 3, 256)), dim3(256), 0, ctx->stream, fg, uj);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_bc_dfield(wm_ctx *ctx, double *df) {
  const FieldGeom &fg = ctx->fg;
  const XSide X = wm_xside(ctx);
  const int cnt = X.xb - X.xa + 1;
  hipLaunchKernelGGL(k_dfield_rows, dim3(wm_cdiv(cnt * 6, 256)), dim3(256), 0, ctx->stream, fg, df,
                     6, X.xa, cnt);
  if (X.wall)
    hipLaunchKernelGGL(k_dfield_walls, dim3(wm_cdiv(fg.gy, 256)), dim3(256), 0, ctx->stream, fg,
                       df, X);
  else
    hipLaunchKernelGGL(k_dfield_cols, dim3(wm_cdiv(fg.gy * 6, 256)), dim3(256), 0, ctx->stream, fg,
                       df, 6);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// phi(nxs-1:nxe+1, nys-1:nye+1) of component l on the device.
int wm_k_bc_phi(wm_ctx *ctx, double *phi, int l) {
  const XSide X = wm_xside(ctx);
  const int nx = X.xb - X.xa + 1, ny = ctx->g.ny;
  hipLaunchKernelGGL(k_phi_rows, dim3(wm_cdiv(nx, 256)), dim3(256), 0, ctx->stream, nx, ny, phi);
  if (X.wall)
    hipLaunchKernelGGL(k_phi_walls, dim3(wm_cdiv(ny + 2, 256)), dim3(256), 0, ctx->stream, nx, ny,
                       phi, l);
  else
    hipLaunchKernelGGL(k_phi_cols, dim3(wm_cdiv(ny + 2, 256)), dim3(256), 0, ctx->stream, nx, ny,
                       phi);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

size_t wm_cg_partials(int nx, int ny) {
  const size_t n = (size_t)nx * ny;
  return 3 * 2 * std::max((size_t)wm_cdiv((long long)n, CGB), (size_t)wm_pc_dot_tiles(nx, ny));
}

// Flexible CG with the learned preconditioner from the warm start in df;
// components that go bad get their warm start back.  its = iterations done.
static int fcg_solve(wm_ctx *ctx, const CgArgs &A, int *its) {
  if (!wm_pc_ready(ctx)) {
    wm_set_error(ctx, WM_ERR_ARG, "flexible CG: the learned preconditioner has no parameters");
    return WM_ERR_ARG;
  }
  const size_t n = A.n;
  if (!ctx->cg_z) {
    WM_HIP(hipMalloc((void **)&ctx->cg_z, sizeof(double) * 3 * n));
    WM_HIP(hipMalloc((void **)&ctx->cg_q, sizeof(double) * 3 * n));
    WM_HIP(hipMalloc((void **)&ctx->cg_x0, sizeof(double) * 3 * n));
    WM_HIP(hipMalloc((void **)&ctx->pc_sc, sizeof(double) * 6));
    WM_HIP(hipMalloc(&ctx->fcg_s, sizeof(FcgScal) * 3));
  }
  FcgScal *S = (FcgScal *)ctx->fcg_s;
  const dim3 grid(A.nblk, 3);
  hipStream_t st = ctx->stream;
  double *pbuf[2] = {ctx->cg_p[0], ctx->cg_p[1]};
  // the warm start, unless the last solve's judge already left it in cg_x0
  if (!ctx->x0_fresh) hipLaunchKernelGGL(k_fcg_save, grid, dim3(CGB), 0, st, A, ctx->df, ctx->cg_x0);
  ctx->x0_fresh = false;
  hipLaunchKernelGGL(k_cg_init, grid, dim3(CGB), 0, st, A, ctx->gkl, ctx->df, ctx->cg_r, pbuf[0],
                     ctx->red);
  hipLaunchKernelGGL(k_fcg_fin, dim3(3), dim3(CGB), 0, st, A, 0, ctx->red, S, ctx->pc_sc,
                     ctx->fcg_max);
  WM_HIP(hipGetLastError());
  // r.z and z.q_prev come out of the preconditioner's last kernel, one
  // partial pair per 16x16 tile
  CgArgs AT = A;
  AT.nblk = wm_pc_dot_tiles(A.nx, A.ny);
  static_assert(3 * sizeof(FcgScal) <= WM_PIN_BYTES, "pinned scratch too small");
  FcgScal *hs = (FcgScal *)ctx->h_pin;
  int cur = 0;
  int batch = std::max(1, ctx->fcg_hint);
  for (;;) {
    WM_HIP(hipMemcpyAsync(hs, S, 3 * sizeof(FcgScal), hipMemcpyDeviceToHost, st));
    WM_HIP(hipStreamSynchronize(st));
    if (!hs[0].active && !hs[1].active && !hs[2].active) break;
    // first the last solve's iteration count, then one per host check;
    // launches after convergence return at once
    for (int bi = 0; bi < batch; ++bi) {
      WM_TRY(wm_pc_apply_dev(ctx, ctx->cg_r, ctx->pc_sc, ctx->cg_z, ctx->cg_q, ctx->red, S));
      hipLaunchKernelGGL(k_fcg_fin, dim3(3), dim3(CGB), 0, st, AT, 1, ctx->red, S, ctx->pc_sc,
                         ctx->fcg_max);
      hipLaunchKernelGGL(k_fcg_dir, grid, dim3(CGB), 0, st, A, ctx->cg_z, pbuf[cur],
                         pbuf[1 - cur], ctx->cg_q, ctx->red, S);
      cur = 1 - cur;
      hipLaunchKernelGGL(k_fcg_fin, dim3(3), dim3(CGB), 0, st, A, 2, ctx->red, S, ctx->pc_sc,
                         ctx->fcg_max);
      hipLaunchKernelGGL(k_fcg_update, grid, dim3(CGB), 0, st, A, ctx->df, ctx->cg_r, pbuf[cur],
                         ctx->cg_q, ctx->red, S);
      hipLaunchKernelGGL(k_fcg_fin, dim3(3), dim3(CGB), 0, st, A, 3, ctx->red, S, ctx->pc_sc,
                         ctx->fcg_max);
      WM_HIP(hipGetLastError());
    }
    batch = 1;
  }
  hipLaunchKernelGGL(k_fcg_restore, grid, dim3(CGB), 0, st, A, ctx->df, ctx->cg_x0, S);
  WM_HIP(hipGetLastError());
  for (int l = 0; l < 3; ++l) its[l] = hs[l].it;
  ctx->fcg_hint = std::max(its[0], std::max(its[1], its[2]));
  return WM_OK;
}

// RHS -> CG -> bc_dfield -> dE -> bc_dfield -> uf += df (field.f90:125-184).
// The current uj must already be folded.
int wm_k_field_solve(wm_ctx *ctx) {
  const wm_grid &g = ctx->g;
  const FieldGeom &fg = ctx->fg;
  const FC f = make_fc(g);
  const int n = g.nx * g.ny;
  const XSide X = wm_xside(ctx);
  hipLaunchKernelGGL(k_rhs, dim3(wm_cdiv(n, 256)), dim3(256), 0, ctx->stream, fg, f, X, ctx->uf,
                     ctx->uj, ctx->gkl);
  WM_HIP(hipGetLastError());

  CgArgs A;
  A.nx = g.nx; A.ny = g.ny; A.gx = fg.gx; A.ps = fg.ps; A.n = (size_t)n;
  A.nbx = wm_cdiv(g.nx, CGB);
  A.nblk = A.nbx * g.ny;
  A.f4 = f.f4; A.f5 = f.f5; A.err = ctx->cg_err; A.ite_max = ctx->ite_max;
  A.X = X;
  const dim3 grid(A.nblk, 3);
  int fcg_it[3] = {0, 0, 0};
  // the learned preconditioner was trained on the periodic operator: the
  // wall families solve with the reference's plain CG
  const bool fcg = ctx->solver == WM_SOLVER_FCG && !X.wall;
  if (fcg) WM_TRY(fcg_solve(ctx, A, fcg_it));
  // the reference's plain CG; after flexible CG it recomputes the true
  // residual and only iterates if that misses the tolerance
  double *pbuf[2] = {ctx->cg_p[0], ctx->cg_p[1]};
  // after flexible CG the judge also copies phi to cg_x0: if no plain
  // iteration follows, that is the next flexible solve's warm start
  double *x0 = fcg ? ctx->cg_x0 : nullptr;
  hipLaunchKernelGGL(k_cg_init, grid, dim3(CGB), 0, ctx->stream, A, ctx->gkl, ctx->df,
                     ctx->cg_r, pbuf[0], ctx->red, x0);
  hipLaunchKernelGGL(k_cg_fin, dim3(3), dim3(CGB), 0, ctx->stream, A, 0, ctx->red, ctx->cg_s);
  WM_HIP(hipGetLastError());
  const bool collect = ctx->solver == WM_SOLVER_CG && !X.wall && wm_pc_collecting(ctx);
  static_assert(3 * sizeof(CgScal) <= WM_PIN_BYTES, "pinned scratch too small");
  CgScal *hs = (CgScal *)ctx->h_pin;
  int iters = 0, cur = 0;
  int batch = collect ? 1 : std::max(1, ctx->cg_hint);
  for (;;) {
    WM_HIP(hipMemcpyAsync(hs, ctx->cg_s, 3 * sizeof(CgScal), hipMemcpyDeviceToHost, ctx->stream));
    WM_HIP(hipStreamSynchronize(ctx->stream));
    bool any = false, all = true;
    for (int l = 0; l < 3; ++l) {
      if (hs[l].err) {
        wm_set_error(ctx, WM_ERR_CG_ITEMAX, "********** stop at cgm after ite_max **********");
        return WM_ERR_CG_ITEMAX;
      }
      any = any || hs[l].active;
      all = all && hs[l].active;
    }
    if (!any) break;
    if (iters >= ctx->ite_max + 1) {
      wm_set_error(ctx, WM_ERR_CG_ITEMAX, "********** stop at cgm after ite_max **********");
      return WM_ERR_CG_ITEMAX;
    }
    // several iterations per host check (first the last plain solve's count,
    // then two): a component that has stopped skips every kernel (its scalars
    // say inactive), so the extra launches change nothing; while pairs are
    // recorded, one iteration at a time
    for (int bi = 0; bi < batch; ++bi) {
      if (collect && all) WM_TRY(wm_pc_snapshot(ctx, iters));
      ++iters;
      hipLaunchKernelGGL(k_cg_pap, grid, dim3(CGB), 0, ctx->stream, A, ctx->cg_r, pbuf[cur],
                         pbuf[1 - cur], ctx->red, ctx->cg_s);
      cur = 1 - cur;
      hipLaunchKernelGGL(k_cg_fin, dim3(3), dim3(CGB), 0, ctx->stream, A, 1, ctx->red, ctx->cg_s);
      hipLaunchKernelGGL(k_cg_update, grid, dim3(CGB), 0, ctx->stream, A, ctx->df, ctx->cg_r,
                         pbuf[cur], ctx->red, ctx->cg_s);
      hipLaunchKernelGGL(k_cg_fin, dim3(3), dim3(CGB), 0, ctx->stream, A, 2, ctx->red, ctx->cg_s);
      WM_HIP(hipGetLastError());
    }
    batch = collect ? 1 : 2;
  }
  if (collect) WM_TRY(wm_pc_snapshot_fin(ctx));
  ctx->x0_fresh = x0 != nullptr && iters == 0;
  if (iters > 0) ctx->cg_hint = std::max(hs[0].ite, std::max(hs[1].ite, hs[2].ite));
  for (int l = 0; l < 3; ++l) ctx->last_iters[l] = fcg_it[l] + hs[l].ite;
  if (fcg) {
    int fb[3] = {hs[0].ite, hs[1].ite, hs[2].ite};
    wm_pc_set_stats(ctx, fcg_it, fb);
  }
  ctx->cg_iters_total += ctx->last_iters[0] + ctx->last_iters[1] + ctx->last_iters[2];

  WM_TRY(wm_k_bc_dfield(ctx, ctx->df));
  hipLaunchKernelGGL(k_delta_e, dim3(wm_cdiv(n, 256)), dim3(256), 0, ctx->stream, fg, f, X,
                     ctx->uf, ctx->uj, ctx->df);
  WM_HIP(hipGetLastError());
  WM_TRY(wm_k_bc_dfield(ctx, ctx->df));
  if (X.xa == 0 && X.xb == g.nx - 1) {
    const size_t nall = 6 * fg.ps;
    hipLaunchKernelGGL(k_add, dim3(wm_cdiv((long long)nall, 256)), dim3(256), 0, ctx->stream, nall,
                       ctx->uf, ctx->df);
  } else {
    const long long nadd = 6LL * fg.gy * (X.xb - X.xa + 5);
    hipLaunchKernelGGL(k_add_cols, dim3(wm_cdiv(nadd, 256)), dim3(256), 0, ctx->stream, fg, X,
                       ctx->uf, ctx->df);
  }
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_energy(wm_ctx *ctx, int bank, double *out) {
  const wm_grid &g = ctx->g;
  const int nseg = g.nsp * g.ny;
  const int nchunk = 8;
  double *part = ctx->diag;
  if (ctx->form == WM_FORM_FP32)   // no settle: a pending mover's slot holds its pushed state
    WM_TRY(wm_k_energy_kin32(ctx, bank, part, nchunk));
  else
    hipLaunchKernelGGL(k_energy_kin, dim3(nchunk, nseg), dim3(256), 0, ctx->stream, g.ny,
                       ctx->cap, g.nsp, ctx->np2, ctx->bank[bank].ux, ctx->bank[bank].uy,
                       ctx->bank[bank].uz, g.c, part, nchunk);
  hipLaunchKernelGGL(k_energy_fld, dim3(g.ny), dim3(256), 0, ctx->stream, ctx->fg, ctx->uf,
                     part + (size_t)nseg * nchunk);
  WM_HIP(hipGetLastError());
  const size_t np_ = (size_t)nseg * nchunk + 2 * (size_t)g.ny;
  double *h = (double *)malloc(sizeof(double) * np_);
  if (!h) return WM_ERR_NOMEM;
  hipError_t e = hipMemcpyAsync(h, part, sizeof(double) * np_, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free(h);
    WM_HIP(e);
  }
  for (int s = 0; s < g.nsp; ++s) {
    double v = 0.0;
    for (int jl = 0; jl < g.ny; ++jl)
      for (int k = 0; k < nchunk; ++k) v += h[((size_t)s * g.ny + jl) * nchunk + k];
    out[s] = g.r[s] * v;
  }
  double b = 0.0, ee = 0.0;
  const double *fp = h + (size_t)nseg * nchunk;
  for (int jl = 0; jl < g.ny; ++jl) {
    b += fp[2 * jl];
    ee += fp[2 * jl + 1];
  }
  out[g.nsp] = ee / (8 * PI);
  out[g.nsp + 1] = b / (8 * PI);
  free(h);
  return WM_OK;
}
