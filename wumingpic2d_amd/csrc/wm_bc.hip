// wm_bc.hip -- particle boundaries: periodic (boundary_periodic.f90:61-248),
// the conducting walls of the shock and reconnection families and the
// shock's injection edge (boundary_shock.f90:62-100, 255-298,
// boundary_reconnection.f90:61-99); their particle_y is the periodic one.
//
// particle_x and the injection edge are elementwise.  particle_y keeps the reference's single-rank
// result exactly, including the order particles end up in inside each row,
// with two kernels over (species, row) segments:
//  pass 1 (one workgroup per segment, in index order): leavers get their y
//   wrapped in round-down arithmetic and are compacted, order preserving,
//   into per-row "up" and "down" mover lists; their slots become the row's
//   ascending hole list (the reference's flag(:,j), :143-169);
//  pass 2 (one workgroup per segment): the arrival list of row j is the
//   concatenation the self-SENDRECVs produce (:174-189) -- for rows other
//   than nys, up-movers of row j-1 then down-movers of row j+1; for row nys,
//   down-movers of row nys+1 then up-movers of row nye -- and the hole
//   filling of :193-236 is evaluated in closed form: hole k takes arrival k;
//   surplus arrivals are appended; with fewer arrivals than holes the row
//   shrinks by m = H-A to L = np2-m, and the k-th remaining hole below L
//   takes the k-th non-hole slot of [L, np2) counted from the top.
#include "wm_internal.h"

namespace {

struct BP {
  int nx, ny, nxgs, nxge, nygs, nyge, nys, nsp, cap, mcap;
  double delx, spanx, spany;
  int xs, xe;          // active columns nxs, nxe (the wall families)
  double wl, wr;       // mirror points 2 (nxs+1) delx and 2 (nxe-1) delx
};

__device__ inline double pred_d(double s) {
  // largest double below s (s finite)
  if (s == 0.0) return -__longlong_as_double(1LL);
  long long b = __double_as_longlong(s);
  b += (s > 0.0) ? -1 : 1;
  return __longlong_as_double(b);
}

// a + b rounded toward -inf (boundary_periodic.f90:74, 124), exact via TwoSum.
__device__ inline double add_rd(double a, double b) {
  const double s = __dadd_rn(a, b);
  const double bb = __dsub_rn(s, a);
  const double err = __dadd_rn(__dsub_rn(a, __dsub_rn(s, bb)), __dsub_rn(b, bb));
  return err < 0.0 ? pred_d(s) : s;
}

__device__ inline int wrap_x(double &x, const BP &P) {
  const int ipos = (int)(x / P.delx);
  if (ipos < P.nxgs) x = add_rd(x, P.spanx);
  else if (ipos >= P.nxge + 1) x = add_rd(x, -P.spanx);
  return ipos;
}

// The wall families' particle_x (boundary_shock.f90:77-89,
// boundary_reconnection.f90:76-88): a particle in a cell below nxs+1 is
// mirrored about (nxs+1) delx, one in a cell at or above nxe-1 about
// (nxe-1) delx, and its momentum reversed.  Returns the new cell.
__device__ inline int wall_x(const Rec64 &R, size_t p, const BP &P) {
  const double x = R.x[p];
  const int ipos = (int)(x / P.delx);
  double w;
  if (ipos < P.xs + 1) w = P.wl;
  else if (ipos >= P.xe - 1) w = P.wr;
  else return ipos;
  const double xn = w - x;
  R.x[p] = xn;
  R.ux[p] = -R.ux[p];
  R.uy[p] = -R.uy[p];
  R.uz[p] = -R.uz[p];
  return (int)(xn / P.delx);
}

__global__ void k_wall_x(BP P, const int *__restrict__ np2, Rec64 B, int *derr) {
  const int seg = blockIdx.y;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  bool bad = false;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int inew = wall_x(B, base + k, P);
    bad = bad || inew < P.nxgs || inew > P.nxge;
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

// boundary_shock__injection (boundary_shock.f90:266-294): the left wall as
// in particle_x; beyond xend (the right edge moved with the upstream flow
// over the step) x mirrors about xend and u_x about u0.
__global__ void k_injection(BP P, const int *__restrict__ np2, Rec64 B, double xend, double u0,
                            int *derr) {
  const int seg = blockIdx.y;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  bool bad = false;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const size_t p = base + k;
    const double x = B.x[p];
    const int ipos = (int)(x / P.delx);
    double xn = x;
    if (ipos < P.xs + 1) {
      xn = P.wl - x;
      B.ux[p] = -B.ux[p];
    } else if (x > xend) {
      xn = 2.0 * xend - x;
      B.ux[p] = 2.0 * u0 - B.ux[p];
    } else {
      continue;
    }
    B.x[p] = xn;
    B.uy[p] = -B.uy[p];
    B.uz[p] = -B.uz[p];
    const int inew = (int)(xn / P.delx);
    bad = bad || inew < P.nxgs || inew > P.nxge;
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

__global__ void k_particle_x(BP P, const int *__restrict__ np2, double *__restrict__ x,
                             int *derr) {
  const int seg = blockIdx.y;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  bool bad = false;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double v = x[base + k];
    const double v0 = v;
    wrap_x(v, P);
    if (v != v0) x[base + k] = v;
    const int inew = (int)(v / P.delx);
    bad = bad || inew < P.nxgs || inew > P.nxge;
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

constexpr int YW = 64;     // one wavefront per segment
constexpr int YT = 4;      // tiles unrolled per walk step

__device__ inline int lanes_below(unsigned long long m) {
  const int lane = threadIdx.x & 63;
  return __popcll(m & ((1ULL << lane) - 1ULL));
}

// Row change of particle p of row j: -1, 0, +1, or 2 (a jump the reference
// does not index).  FP64: jpos = int(y/delx) (boundary_periodic.f90:147),
// with particle_x applied first when fused: xmode 1 the periodic round-down
// wrap, 2 the walls' mirrors.
__device__ inline int row_change(const Rec64 &R, size_t p, int j, int xmode, const BP &P,
                                 int &bad) {
  if (xmode == 2) {
    const int inew = wall_x(R, p, P);
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  } else if (xmode == 1) {
    double xv = R.x[p];
    const double x0 = xv;
    wrap_x(xv, P);
    if (xv != x0) R.x[p] = xv;
    const int inew = (int)(xv / P.delx);
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  }
  const int jpos = (int)(R.y[p] / P.delx);
  const int d = jpos - j;
  return (d >= -1 && d <= 1) ? d : 2;
}

// The record of a leaver as it enters the receiving row: y wraps across the
// periodic boundary in round-down arithmetic (:150-154).
__device__ inline Rec64::Val arriving(const Rec64 &R, size_t p, int j, int d, const BP &P) {
  Rec64::Val v = R.load(p);
  const int jpos = j + d;
  if (jpos <= P.nygs - 1) v.y = add_rd(v.y, P.spany);
  else if (jpos >= P.nyge + 1) v.y = add_rd(v.y, -P.spany);
  return v;
}

// pass 1: one wavefront per segment walks it in index order.
template <class R>
__global__ __launch_bounds__(YW) void k_particle_y1(BP P, int xmode,
                                                    const int *__restrict__ np2, R B,
                                                    const int64_t *__restrict__ id,
                                                    int *__restrict__ hole,
                                                    int *__restrict__ nhole,
                                                    int *__restrict__ nup,
                                                    int *__restrict__ ndn, R mu,
                                                    int64_t *mid_u, R md, int64_t *mid_d,
                                                    int *derr) {
  const int seg = blockIdx.x;
  const int lane = threadIdx.x;
  const int jl = seg % P.ny;
  const int j = P.nys + jl;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  const size_t mb = (size_t)seg * P.mcap;
  int cu = 0, cd = 0;
  int bad = 0;
  for (int k0 = 0; k0 < n; k0 += YW * YT) {
    int dir[YT];
#pragma unroll
    for (int t = 0; t < YT; ++t) {
      const int k = k0 + t * YW + lane;
      dir[t] = k < n ? row_change(B, base + k, j, xmode, P, bad) : 0;
    }
#pragma unroll
    for (int t = 0; t < YT; ++t) {
      const int k = k0 + t * YW + lane;
      if (dir[t] == 2) bad |= DERR_JUMP;
      const bool up = dir[t] == 1, dn = dir[t] == -1;
      const unsigned long long mu_ = __ballot(up), md_ = __ballot(dn);
      if (up || dn) {
        // holes in ascending index order: flag(:,j) of :163-164
        const int hpos = cu + cd + lanes_below(mu_ | md_);
        if (hpos < 2 * P.mcap) hole[2 * mb + hpos] = k;
        else bad |= DERR_MOVER_OVER;
        const int pos = up ? cu + lanes_below(mu_) : cd + lanes_below(md_);
        if (pos < P.mcap) {
          const size_t q = mb + pos;
          (up ? mu : md).store(q, arriving(B, base + k, j, dir[t], P));
          (up ? mid_u : mid_d)[q] = id[base + k];
        } else {
          bad |= DERR_MOVER_OVER;
        }
      }
      cu += __popcll(mu_);
      cd += __popcll(md_);
    }
  }
  if (lane == 0) {
    nhole[seg] = min(cu + cd, 2 * P.mcap);
    nup[seg] = min(cu, P.mcap);
    ndn[seg] = min(cd, P.mcap);
  }
  if (bad) atomicOr(derr, bad);
}

// pass 2: one wavefront per segment.
template <class R>
__global__ __launch_bounds__(YW) void k_particle_y2(BP P, int *__restrict__ np2, R B,
                                                    int64_t *__restrict__ id,
                                                    const int *__restrict__ hole,
                                                    const int *__restrict__ nhole,
                                                    const int *__restrict__ nup,
                                                    const int *__restrict__ ndn, R mu,
                                                    const int64_t *mid_u, R md,
                                                    const int64_t *mid_d, int *derr) {
  const int seg = blockIdx.x;
  const int lane = threadIdx.x;
  const int s = seg / P.ny, jl = seg % P.ny;
  const int below = s * P.ny + (jl == 0 ? P.ny - 1 : jl - 1);
  const int above = s * P.ny + (jl == P.ny - 1 ? 0 : jl + 1);
  // arrival lists, in the order the self-SENDRECVs append them (:174-189):
  // first F, then S
  const bool first_row = (jl == 0);
  const R FB = first_row ? md : mu;
  const int64_t *Fid = first_row ? mid_d : mid_u;
  const int Fseg = first_row ? above : below;
  const int nF = first_row ? ndn[above] : nup[below];
  const R SB = first_row ? mu : md;
  const int64_t *Sid = first_row ? mid_u : mid_d;
  const int Sseg = first_row ? below : above;
  const int nS = first_row ? nup[below] : ndn[above];
  const int A = nF + nS;
  const int H = nhole[seg];
  const int N = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  const int *hl = hole + 2 * (size_t)seg * P.mcap;

  const int newn = N + A - H;
  if (newn > P.cap) {
    if (lane == 0) atomicOr(derr, DERR_NP_OVER);
    return;
  }
  // holes filled by arrivals, and surplus arrivals appended
  for (int a = lane; a < A; a += YW) {
    const int dst = a < H ? hl[a] : N + (a - H);
    const bool f = a < nF;
    const size_t q = (size_t)(f ? Fseg : Sseg) * P.mcap + (f ? a : a - nF);
    B.store(base + dst, (f ? FB : SB).load(q));
    id[base + dst] = (f ? Fid : Sid)[q];
  }
  if (A < H) {
    const int m = H - A;
    int done = 0;    // non-hole tail slots already placed
    for (int q0 = 0; q0 < m; q0 += YW) {
      const int qi = q0 + lane;          // slot N-1-qi, walking down from the top
      bool nonhole = false;
      int src = 0;
      if (qi < m) {
        src = N - 1 - qi;
        // src is a hole iff it is in the sorted list hl[A..H)
        int lo = A, hi = H;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (hl[mid] < src) lo = mid + 1;
          else hi = mid;
        }
        nonhole = !(lo < H && hl[lo] == src);
      }
      const unsigned long long mk = __ballot(nonhole);
      const int nrank = done + lanes_below(mk);
      if (nonhole && A + nrank < H) {
        const int dst = hl[A + nrank];   // the nrank-th remaining hole (below N-m)
        B.store(base + dst, B.load(base + src));
        id[base + dst] = id[base + src];
      }
      done += __popcll(mk);
    }
  }
  if (lane == 0) np2[seg] = newn;
}

BP make_bp(const wm_ctx *ctx) {
  const wm_grid &g = ctx->g;
  BP P;
  P.nx = g.nx; P.ny = g.ny; P.nxgs = g.nxgs; P.nxge = g.nxge; P.nygs = g.nygs;
  P.nyge = g.nyge; P.nys = g.nys; P.nsp = g.nsp; P.cap = ctx->cap; P.mcap = ctx->mcap;
  P.delx = g.delx;
  P.spanx = (double)(g.nxge - g.nxgs + 1) * g.delx;
  P.spany = (double)(g.nyge - g.nygs + 1) * g.delx;
  P.xs = ctx->xe > 0 ? ctx->xs : g.nxgs;
  P.xe = ctx->xe > 0 ? ctx->xe : g.nxge;
  P.wl = 2.0 * (P.xs + 1) * g.delx;
  P.wr = 2.0 * (P.xe - 1) * g.delx;
  return P;
}

}  // namespace

int wm_k_particle_x(wm_ctx *ctx, int bank) {
  BP P = make_bp(ctx);
  const int nseg = ctx->g.nsp * ctx->g.ny;
  const int nb = std::min(64, std::max(1, wm_cdiv(ctx->cap, 256)));
  if (ctx->bcf != WM_BC_PERIODIC)
    hipLaunchKernelGGL(k_wall_x, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->np2,
                       ctx->bank[bank], ctx->derr);
  else
    hipLaunchKernelGGL(k_particle_x, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->np2,
                       ctx->bank[bank].x, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// xend of boundary_shock__injection (boundary_shock.f90:269), as the
// reference evaluates it.
double wm_injection_xend(const wm_ctx *ctx, double u0) {
  const wm_grid &g = ctx->g;
  const int nxe = ctx->xe > 0 ? ctx->xe : g.nxge;
  return nxe * g.delx + u0 / sqrt(1 + (u0 * u0) / (g.c * g.c)) * g.delt;
}

int wm_k_injection(wm_ctx *ctx, int bank, double u0) {
  BP P = make_bp(ctx);
  const int nseg = ctx->g.nsp * ctx->g.ny;
  const int nb = std::min(64, std::max(1, wm_cdiv(ctx->cap, 256)));
  hipLaunchKernelGGL(k_injection, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->np2,
                     ctx->bank[bank], wm_injection_xend(ctx, u0), u0, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_particle_y(wm_ctx *ctx, int bank, int xmode) {
  BP P = make_bp(ctx);
  const int nseg = ctx->g.nsp * ctx->g.ny;
  int64_t *id = ctx->id[ctx->idcur];
  if (ctx->form == WM_FORM_FP32) {
    // the FP32 form migrates particles between its cell bins (wm_sort.hip)
    wm_set_error(ctx, WM_ERR_ARG, "particle_y works on the FP64 particle form");
    return WM_ERR_ARG;
  }
  {
    const Rec64 &B = ctx->bank[bank];
    hipLaunchKernelGGL(k_particle_y1<Rec64>, dim3(nseg), dim3(YW), 0, ctx->stream, P, xmode,
                       ctx->np2, B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, ctx->derr);
    hipLaunchKernelGGL(k_particle_y2<Rec64>, dim3(nseg), dim3(YW), 0, ctx->stream, P, ctx->np2,
                       B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  return WM_OK;
}
