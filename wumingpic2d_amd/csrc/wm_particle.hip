// wm_particle.hip -- Buneman-Boris push and Esirkepov deposit (FP64 form).
//
// Work decomposition (the same for the push, the deposit and the fused
// push+deposit kernel): one workgroup per (row j, chunk of CHUNK cells); the
// workgroup's 256 lanes form 32 groups of 8 lanes, group g owns cell
// i0+g and its 8 lanes stride through that cell's particles of every species
// (the reference's cell -> species -> particle loop, particle.f90:85-94 and
// field.f90:212-221).  Particles of one cell are contiguous (sorted); a
// group's windows of 16 particles start at the 16-particle boundary at or
// below the cell's first particle (lanes outside the cell idle), and lane
// sub takes the window's particles sub and 8 + sub, so the group's two
// loads of an array cover one aligned 128-byte line.
//
// Push fields come from a per-workgroup LDS tile of the cell-centred fields.
// Deposit: each lane accumulates the stay-in-cell stencils of its particles
// in registers (all in the same cell), the 8 lanes of a group combine by
// butterfly shuffles; cell-crossing particles are listed in LDS as they are
// found and their full 5x5 stencils summed after the cell's loop.  The
// workgroup assembles its 5 x (CHUNK+4) current tile in LDS in a fixed
// order.  Tiles go to a partial buffer that a gather
// kernel sums in a fixed order, so the deposit is deterministic (no
// floating-point atomics).
#include "wm_internal.h"

namespace {

constexpr int CHUNK = 32;             // cells per workgroup
constexpr int GROUP = 8;              // lanes per cell
constexpr int TW = CHUNK + 2;         // push tile width (halo 1)
constexpr int DW = CHUNK + 4;         // deposit tile width (halo 2)
constexpr int TILE = 5 * DW * 3;      // doubles per deposit tile
constexpr int NACC = 65;              // Jx 4x5 + Jy 5x4 + Jz 5x5
constexpr int MCAP = 16;              // a cell's movers listed in LDS per species

struct PP {                           // particle-kernel parameters
  int nx, ny, nxgs, nys, nsp, cap, nchunk;
  double d_delx, delt, c, qc_scale;   // qc_scale = delx / delt
  double q[WM_MAXSP], r[WM_MAXSP];
  // the shock's injection edge, applied to the pushed particle before its
  // current (proj/shock/app.f90:112-114): inj, nxs, the left mirror point
  // 2 (nxs+1) delx, xend and u0 (boundary_shock.f90:266-294)
  int inj, xs;
  double delx, wl, xend, u0;
};

__device__ inline void injection_edge(double &x, double &ux, double &uy, double &uz,
                                      const PP &P) {
  if ((int)(x / P.delx) < P.xs + 1) {
    x = P.wl - x;
    ux = -ux;
  } else if (x > P.xend) {
    x = 2.0 * P.xend - x;
    ux = 2.0 * P.u0 - ux;
  } else {
    return;
  }
  uy = -uy;
  uz = -uz;
}

__device__ inline void shape(double dh, double &sm, double &s0, double &sp) {
  // 2nd-order (TSC) weights: particle.f90:98-100
  sm = 0.5 * (0.5 - dh) * (0.5 - dh);
  s0 = 0.75 - dh * dh;
  sp = 0.5 * (0.5 + dh) * (0.5 + dh);
}

// Cell-centred field tile: T[(k*3 + b)*TW + col], b = row j-1..j+1,
// col = i - (i0 - 1).
__device__ inline void load_tmp_tile(double *T, const double *tmp, int TX,
                                     size_t tps, int jl, int c0, int ncol) {
  // tmp planes hold cells (nxs-1 .. nxe+1) x (nys-1 .. nye+1); local column
  // index of cell i is i - (nxs-1); c0 = local column of i0 - 1.
  for (int t = threadIdx.x; t < 6 * 3 * TW; t += blockDim.x) {
    const int col = t % TW;
    const int b = (t / TW) % 3;
    const int k = t / (3 * TW);
    double v = 0.0;
    if (col < ncol) v = tmp[(size_t)k * tps + (size_t)(c0 + col) + (size_t)TX * (size_t)(jl + b)];
    T[t] = v;
  }
}

struct PushOut { double x, y, ux, uy, uz, ig; };   // ig = 1 / gamma of the new u

__device__ inline PushOut boris(const double *T, int gcol, double x, double y,
                                double ux, double uy, double uz, int i, int j,
                                const PP &P, double fac1, double txxx, double fac2) {
  // particle.f90:96-161
  double sx0, sx1, sx2, sy0, sy1, sy2;
  shape(x * P.d_delx - 0.5 - i, sx0, sx1, sx2);
  shape(y * P.d_delx - 0.5 - j, sy0, sy1, sy2);
  double f[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double *Tk = T + k * 3 * TW + gcol;   // gcol = column of i-1
    f[k] = (Tk[0] * sx0 + Tk[1] * sx1 + Tk[2] * sx2) * sy0 +
           (Tk[TW] * sx0 + Tk[TW + 1] * sx1 + Tk[TW + 2] * sx2) * sy1 +
           (Tk[2 * TW] * sx0 + Tk[2 * TW + 1] * sx1 + Tk[2 * TW + 2] * sx2) * sy2;
  }
  const double bpx = f[0], bpy = f[1], bpz = f[2];
  const double epx = f[3], epy = f[4], epz = f[5];
  double uvm1 = ux + fac1 * epx;
  double uvm2 = uy + fac1 * epy;
  double uvm3 = uz + fac1 * epz;
  const double gam = sqrt(P.c * P.c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
  const double igam = 1.0 / gam;
  const double fac1r = fac1 * igam;
  const double fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
  const double uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
  const double uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
  const double uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
  uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
  uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
  uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
  PushOut o;
  o.ux = uvm1 + fac1 * epx;
  o.uy = uvm2 + fac1 * epy;
  o.uz = uvm3 + fac1 * epz;
  const double ig = 1.0 / sqrt(1.0 + (o.ux * o.ux + o.uy * o.uy + o.uz * o.uz) / (P.c * P.c));
  o.x = x + o.ux * P.delt * ig;
  o.y = y + o.uy * P.delt * ig;
  o.ig = ig;
  return o;
}

// Shape data of one particle's move for the Esirkepov deposit
// (field.f90:224-268): s0 at the old position around cell (i, j), ds = s1
// shifted by the cell increment minus s0.  Returns false when the particle
// jumped more than one cell.
struct Move {
  double s0x[5], s0y[5], dsx[5], dsy[5];
};

__device__ inline void place_s1(double a, double b, double c, int inc, double *ds) {
  // s1 placed at offset inc (the smo_1..3 selection of field.f90:244-251)
  ds[0] = inc < 0 ? a : 0.0;
  ds[1] = inc < 0 ? b : (inc == 0 ? a : 0.0);
  ds[2] = inc < 0 ? c : (inc == 0 ? b : a);
  ds[3] = inc > 0 ? b : (inc == 0 ? c : 0.0);
  ds[4] = inc > 0 ? c : 0.0;
}

__device__ inline bool make_move(Move &m, double x0, double y0, double x1, double y1, int i,
                                 int j, const PP &P) {
  m.s0x[0] = m.s0x[4] = m.s0y[0] = m.s0y[4] = 0.0;
  shape(x0 * P.d_delx - 0.5 - i, m.s0x[1], m.s0x[2], m.s0x[3]);
  shape(y0 * P.d_delx - 0.5 - j, m.s0y[1], m.s0y[2], m.s0y[3]);
  const int i2 = (int)(x1 * P.d_delx), j2 = (int)(y1 * P.d_delx);
  double a, b, c;
  shape(x1 * P.d_delx - 0.5 - i2, a, b, c);
  place_s1(a, b, c, i2 - i, m.dsx);
  shape(y1 * P.d_delx - 0.5 - j2, a, b, c);
  place_s1(a, b, c, j2 - j, m.dsy);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    m.dsx[k] -= m.s0x[k];
    m.dsy[k] -= m.s0y[k];
  }
  return (i2 - i >= -1 && i2 - i <= 1) && (j2 - j >= -1 && j2 - j <= 1);
}

// Full 5x5 stencil of one current component (field.f90:274-298) into acc:
//  C = 0: Jx, acc[a*5+b], tile column a+1 (a = 0..3), pjx(ip+1,jp) =
//         pjx(ip,jp) - q dx/dt ds_x(ip) (s0_y(jp) + ds_y(jp)/2)
//  C = 1: Jy, acc[a*4+b], tile row b+1 (b = 0..3), the same along y
//  C = 2: Jz, acc[a*5+b], q gvz (s0x s0y + dsx s0y/2 + s0x dsy/2 + dsx dsy/3)
template <int C>
__device__ inline void stencil(double *acc, const Move &m, double qc, double qv) {
  if (C == 0) {
    double w = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      w -= qc * m.dsx[a];
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[a * 5 + b] += w * (m.s0y[b] + 0.5 * m.dsy[b]);
    }
  } else if (C == 1) {
    double w = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      w -= qc * m.dsy[b];
#pragma unroll
      for (int a = 0; a < 5; ++a) acc[a * 4 + b] += (m.s0x[a] + 0.5 * m.dsx[a]) * w;
    }
  } else {
    const double third = 1.0 / 3.0;
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      const double za = qv * (m.s0x[a] + 0.5 * m.dsx[a]);
      const double zb = qv * (0.5 * m.s0x[a] + third * m.dsx[a]);
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[a * 5 + b] += za * m.s0y[b] + zb * m.dsy[b];
    }
  }
}

// A particle that stays in its cell has s0 and ds supported on offsets 1..3
// only, so of the 65 stencil values just the 27 below can be non-zero
// (every other value is an exact zero in the reference's arithmetic):
//   Jx tile columns 2..4 x rows 1..3, Jy columns 1..3 x rows 2..4,
//   Jz columns 1..3 x rows 1..3.
// core[0..8] Jx (a-1)*3+(b-1) with a = W index 1..3; core[9..17] Jy
// (a-1)*3+(b-1) with b = W index 1..3; core[18..26] Jz (a-1)*3+(b-1).
__device__ inline void stencil_core(double *core, double x0, double y0, double x1,
                                    double y1, double gvz, int i, int j, double q,
                                    const PP &P) {
  double sx[3], sy[3], dx[3], dy[3];
  shape(x0 * P.d_delx - 0.5 - i, sx[0], sx[1], sx[2]);
  shape(y0 * P.d_delx - 0.5 - j, sy[0], sy[1], sy[2]);
  shape(x1 * P.d_delx - 0.5 - i, dx[0], dx[1], dx[2]);
  shape(y1 * P.d_delx - 0.5 - j, dy[0], dy[1], dy[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dx[k] -= sx[k];
    dy[k] -= sy[k];
  }
  const double qc = q * P.qc_scale;
  const double qv = q * gvz;
  double X[3], Y[3], wx[3], wy[3];
  double w1 = 0.0, w2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] = sx[k] + 0.5 * dx[k];
    Y[k] = sy[k] + 0.5 * dy[k];
    w1 -= qc * dx[k];
    w2 -= qc * dy[k];
    wx[k] = w1;
    wy[k] = w2;
  }
  const double third = 1.0 / 3.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double za = qv * X[a];
    const double zb = qv * (0.5 * sx[a] + third * dx[a]);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      core[a * 3 + b] += wx[a] * Y[b];
      core[9 + a * 3 + b] += X[a] * wy[b];
      core[18 + a * 3 + b] += za * sy[b] + zb * dy[b];
    }
  }
}

// Index of a core value in the 65-value cell layout of acc_at().
__device__ inline int core_slot(int c) {
  const int comp = c / 9, a = (c % 9) / 3 + 1, b = c % 3 + 1;
  if (comp == 0) return a * 5 + b;             // Jx: W index a -> acc (a, b)
  if (comp == 1) return 20 + a * 4 + b;        // Jy: acc (a, W index b)
  return 40 + a * 5 + b;
}

// Value of accumulator set `cb` (one cell) at tile offset (a, b) for comp.
__device__ inline double acc_at(const double *cb, int comp, int a, int b) {
  if (comp == 0) return a >= 1 ? cb[(a - 1) * 5 + b] : 0.0;
  if (comp == 1) return b >= 1 ? cb[20 + a * 4 + (b - 1)] : 0.0;
  return cb[40 + a * 5 + b];
}

__device__ inline double group_sum(double v) {
  v += __shfl_xor(v, 1, GROUP);
  v += __shfl_xor(v, 2, GROUP);
  v += __shfl_xor(v, 4, GROUP);
  return v;
}

// Contribution of one component of a cell's crossing particles, combined
// over the group and added by lane 0 to cellbuf.  The cell's first MCAP
// movers of the species come from its LDS list L (x0, y0, x1, y1, q v_z;
// lane sub takes entries sub, sub + GROUP, ...); a mover past them from the
// lane's own mask (bit m = the lane's m-th particle from ka, the state read
// back from the banks); particles beyond the mask are re-examined.
template <int C>
__device__ inline void crossing_pass(double *cb, int sub, const double *L, int nl,
                                     unsigned long long mask, int ka, int ke, size_t base,
                                     const double *xa, const double *ya, const double *xb,
                                     const double *yb, const double *uxb, const double *uyb,
                                     const double *uzb, int i, int j, double q, const PP &P) {
  constexpr int N = C == 2 ? 25 : 20;
  double acc[N];
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = 0.0;
  const double qc = q * P.qc_scale;
  for (int e = sub; e < nl; e += GROUP) {
    Move m;
    make_move(m, L[e], L[MCAP + e], L[2 * MCAP + e], L[3 * MCAP + e], i, j, P);
    stencil<C>(acc, m, qc, L[4 * MCAP + e]);
  }
  auto one = [&](int k, bool known) {
    const size_t p = base + (size_t)k;
    Move m;
    const double x1 = xb[p], y1 = yb[p];
    const bool cross = (int)(x1 * P.d_delx) != i || (int)(y1 * P.d_delx) != j;
    if (!known && !cross) return;
    make_move(m, xa[p], ya[p], x1, y1, i, j, P);
    double qv = 0.0;
    if (C == 2) {
      const double ux = uxb[p], uy = uyb[p], uz = uzb[p];
      qv = q * (uz / sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (P.c * P.c)));
    }
    stencil<C>(acc, m, qc, qv);
  };
  while (mask) {
    const int m = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    one(ka + sub + GROUP * m, true);
  }
  for (int k = ka + sub + GROUP * 64; k < ke; k += GROUP) one(k, false);
  const int off = C == 0 ? 0 : (C == 1 ? 20 : 40);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double v = group_sum(acc[k]);
    if (sub == 0) cb[off + k] += v;
  }
}

// Element k of a row segment that starts at seg (the same in every lane):
// a 32-bit byte offset from a scalar base, so that no per-lane 64-bit
// address is kept per array.  A segment holds fewer than 2^29 doubles.
__device__ inline double seg_ld(const double *seg, int k) {
  return *(const double *)((const char *)seg + (unsigned)k * 8u);
}
__device__ inline void seg_st(double *seg, int k, double v) {
  *(double *)((char *)seg + (unsigned)k * 8u) = v;
}

template <bool PUSH, bool DEPOSIT>
__global__ __launch_bounds__(256, 3) void k_particle(
    PP P, const int *__restrict__ cumcnt,
    const double *__restrict__ tmp, int TX, size_t tps,
    const double *__restrict__ xa, const double *__restrict__ ya,
    const double *__restrict__ uxa, const double *__restrict__ uya,
    const double *__restrict__ uza,
    double *__restrict__ xb, double *__restrict__ yb,
    double *__restrict__ uxb, double *__restrict__ uyb,
    double *__restrict__ uzb,
    double *__restrict__ part, int *__restrict__ derr) {
  __shared__ double T[PUSH ? 6 * 3 * TW : 1];
  __shared__ double cellbuf[DEPOSIT ? CHUNK * NACC : 1];
  // per cell: its first MCAP movers of the species in hand, as
  // [x0 | y0 | x1 | y1 | q v_z][entry], so that their full stencils need not
  // read the banks again (their lines are gone from L2 by then)
  __shared__ double mvl[DEPOSIT ? CHUNK * 5 * MCAP : 1];
  const int chunk = blockIdx.x;
  const int jl = blockIdx.y;                 // row j = nys + jl
  const int j = P.nys + jl;
  const int i0 = P.nxgs + chunk * CHUNK;
  const int ncell = min(CHUNK, P.nx - chunk * CHUNK);
  if (PUSH) load_tmp_tile(T, tmp, TX, tps, jl, chunk * CHUNK, min(TW, ncell + 2));
  if (DEPOSIT)
    for (int e = threadIdx.x; e < CHUNK * NACC; e += blockDim.x) cellbuf[e] = 0.0;
  __syncthreads();
  const int g = threadIdx.x / GROUP;
  const int sub = threadIdx.x % GROUP;
  const int i = i0 + g;
  double *cb = cellbuf + g * NACC;
  double *L = mvl + (DEPOSIT ? g * 5 * MCAP : 0);
  const int wg = (threadIdx.x & 63) / GROUP;   // the group's place in its wavefront
  double core[DEPOSIT ? 27 : 1];
  if (DEPOSIT) {
#pragma unroll
    for (int k = 0; k < 27; ++k) core[k] = 0.0;
  }
  bool ok = true;
  if (g < ncell) {
    for (int s = 0; s < P.nsp; ++s) {
      const size_t seg = (size_t)s * P.ny + jl;
      const int *cc = cumcnt + seg * (size_t)(P.nx + 1);
      const int kb = max(0, cc[i - P.nxgs]), ke = min(P.cap, cc[i - P.nxgs + 1]);
      const size_t base = seg * (size_t)P.cap;
      const double q = P.q[s], r = P.r[s];
      const double fac1 = q / r * 0.5 * P.delt;
      const double txxx = fac1 * fac1;
      const double fac2 = q * P.delt / r;
      unsigned long long mask = 0ULL;
      if (kb >= ke) continue;                  // no particle of species s here
      // windows of 2 GROUP consecutive particles start at the 2 GROUP
      // boundary at or below the cell's first particle, so that each array's
      // load is one aligned 128-byte line for the group; lane sub takes the
      // window's particles sub and GROUP + sub, in that order (the order of
      // GROUP-particle windows, m counting them from ka)
      const int ka = kb - kb % (2 * GROUP);
      // segment bases: the same for the whole workgroup, 32-bit indices below
      const double *Xa = xa + base, *Ya = ya + base;
      const double *Uxa = uxa + base, *Uya = uya + base, *Uza = uza + base;
      double *Xb = xb + base, *Yb = yb + base;
      double *Uxb = uxb + base, *Uyb = uyb + base, *Uzb = uzb + base;
      int mcnt = 0;                            // the cell's movers so far (lane 0's is kept)
      // particle k (the lane's m-th): pushed and stored, its stay-in-cell
      // stencil added; a crossing one returns true with its deposit inputs
      auto one = [&](int k, int m, double x0, double y0, double cux, double cuy, double cuz,
                     double &x1, double &y1, double &qv) {
        if (k < kb || k >= ke) return false;   // before or after the cell
        double ux, uy, uz, ig = 0.0;
        // keep the 54 tile values in LDS instead of hoisting them into
        // registers for the whole cell loop (occupancy)
        __asm__ volatile("" ::: "memory");
        if (PUSH) {
          PushOut o = boris(T, g, x0, y0, cux, cuy, cuz, i, j, P, fac1, txxx, fac2);
          if (P.inj) injection_edge(o.x, o.ux, o.uy, o.uz, P);
          seg_st(Xb, k, o.x); seg_st(Yb, k, o.y);
          seg_st(Uxb, k, o.ux); seg_st(Uyb, k, o.uy); seg_st(Uzb, k, o.uz);
          x1 = o.x; y1 = o.y; ux = o.ux; uy = o.uy; uz = o.uz; ig = o.ig;
        } else {
          x1 = seg_ld(Xb, k); y1 = seg_ld(Yb, k);
          ux = seg_ld(Uxb, k); uy = seg_ld(Uyb, k); uz = seg_ld(Uzb, k);
        }
        if (DEPOSIT) {
          const int i2 = (int)(x1 * P.d_delx), j2 = (int)(y1 * P.d_delx);
          // field.f90:270's v_z = u_z / gamma; the fused kernel reuses the
          // push's 1 / gamma of the same u (a product instead of a
          // quotient: within an ulp), except where the shock's injection
          // edge may have reflected the pushed u
          const double gvz = PUSH && !P.inj
                                 ? uz * ig
                                 : uz / sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (P.c * P.c));
          if (i2 == i && j2 == j) {
            stencil_core(core, x0, y0, x1, y1, gvz, i, j, q, P);
          } else {
            ok = ok && (i2 - i >= -1 && i2 - i <= 1 && j2 - j >= -1 && j2 - j <= 1);
            qv = q * gvz;
            return m < 64;                     // later ones: re-examined from the banks
          }
        }
        return false;
      };
      // a crossing particle goes to the cell's LDS list in (window, lane)
      // order, ranked by a ballot over the group (lane 0 of the group runs
      // every window any of its lanes runs); past MCAP to the lane's mask
      auto record = [&](bool cross, int m, double x0, double y0, double x1, double y1,
                        double qv) {
        const unsigned gb = (unsigned)(__ballot(cross) >> (GROUP * wg)) & 0xFFu;
        if (cross) {
          const int e = mcnt + __popc(gb & ((1u << sub) - 1u));
          if (e < MCAP) {
            L[e] = x0;
            L[MCAP + e] = y0;
            L[2 * MCAP + e] = x1;
            L[3 * MCAP + e] = y1;
            L[4 * MCAP + e] = qv;
          } else {
            mask |= 1ULL << m;
          }
        }
        mcnt += __popc(gb);
      };
      for (int m = 0, k = ka + sub; k < ke; k += 2 * GROUP, m += 2) {
        // both halves' loads issued together, each index clamped into the
        // cell's particles (a lane outside them reads a neighbour and idles)
        const int kA = min(max(k, kb), ke - 1), kB = min(max(k + GROUP, kb), ke - 1);
        const double xA = seg_ld(Xa, kA), yA = seg_ld(Ya, kA);
        const double xB = seg_ld(Xa, kB), yB = seg_ld(Ya, kB);
        double uxA = 0, uyA = 0, uzA = 0, uxB = 0, uyB = 0, uzB = 0;
        if (PUSH) {
          uxA = seg_ld(Uxa, kA); uyA = seg_ld(Uya, kA); uzA = seg_ld(Uza, kA);
          uxB = seg_ld(Uxa, kB); uyB = seg_ld(Uya, kB); uzB = seg_ld(Uza, kB);
        }
        // one particle at a time: interleaving the two would need twice the
        // registers of one push and deposit
        double x1, y1, qv = 0.0;
        bool cr = one(k, m, xA, yA, uxA, uyA, uzA, x1, y1, qv);
        if (DEPOSIT) record(cr, m, xA, yA, x1, y1, qv);
        __builtin_amdgcn_sched_barrier(0);
        cr = one(k + GROUP, m + 1, xB, yB, uxB, uyB, uzB, x1, y1, qv);
        if (DEPOSIT) record(cr, m + 1, xB, yB, x1, y1, qv);
      }
      if (DEPOSIT) {
        // add the stay-in-cell stencils of this species over the 8 lanes
#pragma unroll
        for (int c = 0; c < 27; ++c) {
          const double v = group_sum(core[c]);
          core[c] = 0.0;
          if (sub == 0) cb[core_slot(c)] += v;
        }
      }
      mcnt = __shfl(mcnt, 0, GROUP);
      __builtin_amdgcn_wave_barrier();         // the group reads what its lanes listed
      const int nl = DEPOSIT ? min(mcnt, MCAP) : 0;
      if (DEPOSIT && __ballot(nl > 0 || mask != 0ULL || ke - ka > GROUP * 64)) {
        // cell-crossing particles: full 5x5 stencils, one component at a time
        // (the banks' new state stream-ordered after this lane's own writes)
        crossing_pass<0>(cb, sub, L, nl, mask, ka, ke, base, xa, ya, xb, yb, uxb, uyb, uzb, i, j,
                         q, P);
        crossing_pass<1>(cb, sub, L, nl, mask, ka, ke, base, xa, ya, xb, yb, uxb, uyb, uzb, i, j,
                         q, P);
        crossing_pass<2>(cb, sub, L, nl, mask, ka, ke, base, xa, ya, xb, yb, uxb, uyb, uzb, i, j,
                         q, P);
      }
      __builtin_amdgcn_wave_barrier();         // the list is read before the next species
    }
  }
  if (!ok) atomicOr(derr, DERR_JUMP);
  if (DEPOSIT) {
    __syncthreads();
    // assemble the workgroup tile [b][t][comp] in a fixed order
    double *out = part + ((size_t)jl * P.nchunk + chunk) * TILE;
    for (int e = threadIdx.x; e < TILE; e += blockDim.x) {
      const int comp = e % 3;
      const int t = (e / 3) % DW;
      const int b = e / (3 * DW);
      double sum = 0.0;
      const int glo = max(0, t - 4), ghi = min(ncell - 1, t);
      for (int gg = glo; gg <= ghi; ++gg) sum += acc_at(cellbuf + gg * NACC, comp, t - gg, b);
      out[e] = sum;
    }
  }
}

// Sum of the deposit tiles into uj (ghost-inclusive), fixed order:
// source row offset b ascending, chunk ascending.
__global__ void k_deposit_gather(PP P, FieldGeom fg, const double *__restrict__ part,
                                 double *__restrict__ uj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = fg.gx * fg.gy;
  if (t >= n) return;
  const int li = t % fg.gx, lj = t / fg.gx;     // ghost-inclusive local indices
  // local column li <-> reference i = nxgs - 2 + li; tile column of chunk c:
  // tcol = i - (nxgs + CHUNK*c) + 2 = li - CHUNK*c
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int chi = min(li / CHUNK, P.nchunk - 1);
  for (int b = 0; b < 5; ++b) {
    // source row jl with jl + (b - 2) = lj - 2  ->  jl = lj - b
    const int jl = lj - b;
    if (jl < 0 || jl >= P.ny) continue;
    for (int c = max(0, chi - 1); c <= chi; ++c) {
      const int tc = li - CHUNK * c;
      if (tc < 0 || tc >= DW) continue;
      const double *tp = part + ((size_t)jl * P.nchunk + c) * TILE + ((size_t)b * DW + tc) * 3;
      s0 += tp[0];
      s1 += tp[1];
      s2 += tp[2];
    }
  }
  uj[t] = s0;
  uj[fg.ps + t] = s1;
  uj[2 * fg.ps + t] = s2;
}

// Cell-centred fields, particle.f90:68-81, on cells nxs-1..nxe+1, nys-1..nye+1.
__global__ void k_cell_fields(FieldGeom fg, const double *__restrict__ uf,
                              double *__restrict__ tmp, int TX, int TY, size_t tps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= TX * TY) return;
  const int a = t % TX, b = t / TX;
  // cell (i, j) = (nxs-1+a, nys-1+b) -> ghost-inclusive local (a+1, b+1)
  const size_t c0 = (size_t)(a + 1) + (size_t)fg.gx * (size_t)(b + 1);
  const size_t gx = fg.gx;
  const double *B1 = uf, *B2 = uf + fg.ps, *B3 = uf + 2 * fg.ps;
  const double *E1 = uf + 3 * fg.ps, *E2 = uf + 4 * fg.ps, *E3 = uf + 5 * fg.ps;
  tmp[t] = 0.5 * (B1[c0] + B1[c0 + gx]);
  tmp[tps + t] = 0.5 * (B2[c0] + B2[c0 + 1]);
  tmp[2 * tps + t] = 0.25 * (B3[c0] + B3[c0 + 1] + B3[c0 + gx] + B3[c0 + gx + 1]);
  tmp[3 * tps + t] = 0.5 * (E1[c0] + E1[c0 + 1]);
  tmp[4 * tps + t] = 0.5 * (E2[c0] + E2[c0 + gx]);
  tmp[5 * tps + t] = E3[c0];
}

PP make_pp(const wm_ctx *ctx) {
  PP P;
  const wm_grid &g = ctx->g;
  P.nx = g.nx; P.ny = g.ny; P.nxgs = g.nxgs; P.nys = g.nys; P.nsp = g.nsp;
  P.cap = ctx->cap;
  P.nchunk = wm_cdiv(g.nx, CHUNK);
  P.d_delx = 1.0 / g.delx;
  P.delt = g.delt;
  P.c = g.c;
  P.qc_scale = g.delx * (1.0 / g.delt);
  for (int s = 0; s < WM_MAXSP; ++s) {
    P.q[s] = s < g.nsp ? g.q[s] : 0.0;
    P.r[s] = s < g.nsp ? g.r[s] : 1.0;
  }
  P.inj = 0;
  P.xs = ctx->xe > 0 ? ctx->xs : g.nxgs;
  P.delx = g.delx;
  P.wl = 2.0 * (P.xs + 1) * g.delx;
  P.xend = 0.0;
  P.u0 = 0.0;
  return P;
}

template <bool PUSH, bool DEP>
int launch_particle(wm_ctx *ctx, int a, int b, bool inj = false) {
  PP P = make_pp(ctx);
  if (inj) {
    P.inj = 1;
    P.u0 = ctx->u0;
    P.xend = wm_injection_xend(ctx, ctx->u0);
  }
  dim3 grid(P.nchunk, P.ny);
  const int TX = ctx->g.nx + 2;
  const size_t tps = (size_t)TX * (size_t)(ctx->g.ny + 2);
  const PBank &A = ctx->bank[a];
  const PBank &B = ctx->bank[b];
  hipLaunchKernelGGL((k_particle<PUSH, DEP>), grid, dim3(256), 0, ctx->stream, P,
                     ctx->cumcnt, ctx->tmp, TX, tps, A.x, A.y, A.ux, A.uy, A.uz,
                     B.x, B.y, B.ux, B.uy, B.uz, ctx->dep_part, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

}  // namespace

size_t wm_deposit_part_elems(int nx, int ny) {
  return (size_t)ny * (size_t)wm_cdiv(nx, CHUNK) * (size_t)TILE;
}

int wm_k_cell_fields(wm_ctx *ctx) {
  const int TX = ctx->g.nx + 2, TY = ctx->g.ny + 2;
  const size_t tps = (size_t)TX * TY;
  hipLaunchKernelGGL(k_cell_fields, dim3(wm_cdiv((long long)TX * TY, 256)), dim3(256), 0,
                     ctx->stream, ctx->fg, ctx->uf, ctx->tmp, TX, TY, tps);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_push(wm_ctx *ctx, int src, int dst) { return launch_particle<true, false>(ctx, src, dst); }
int wm_k_deposit(wm_ctx *ctx, int oldb, int newb) { return launch_particle<false, true>(ctx, oldb, newb); }
// The device-resident shock step fuses its injection edge between the push
// and the current.
int wm_k_push_deposit(wm_ctx *ctx, int src, int dst) {
  return launch_particle<true, true>(ctx, src, dst, ctx->bcf == WM_BC_SHOCK);
}

int wm_k_deposit_gather(wm_ctx *ctx) {
  PP P = make_pp(ctx);
  const int n = ctx->fg.gx * ctx->fg.gy;
  hipLaunchKernelGGL(k_deposit_gather, dim3(wm_cdiv(n, 256)), dim3(256), 0, ctx->stream, P,
                     ctx->fg, ctx->dep_part, ctx->uj);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
