// wm_bc.hip -- particle boundaries: periodic (boundary_periodic.f90:61-248),
// the conducting walls of the shock and reconnection families and the
// shock's injection edge (boundary_shock.f90:62-100, 255-298,
// boundary_reconnection.f90:61-99); their particle_y is the periodic one.
//
// particle_x and the injection edge are elementwise.  particle_y keeps the reference's single-rank
// result exactly, including the order particles end up in inside each row,
// with two kernels over (species, row) segments:
//  pass 1 (one workgroup of YNW wavefronts per segment): leavers get their y
//   wrapped in round-down arithmetic and are compacted, order preserving,
//   into per-row "up" and "down" mover lists; their slots become the row's
//   ascending hole list (the reference's flag(:,j), :143-169).  The segment
//   is taken YST tiles of 64 particles at a time: the wavefronts classify
//   the tiles (ballot masks in LDS), the workgroup scans the tiles' mover
//   counts, and each tile's movers then go to their places in index order;
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
__device__ inline int wall_x_at(const Rec64 &R, size_t p, double x, const BP &P) {
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
__device__ inline int wall_x(const Rec64 &R, size_t p, const BP &P) { return wall_x_at(R, p, R.x[p], P); }

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

constexpr int YW = 64;     // wavefront
constexpr int YNW = 8;     // wavefronts per segment (pass 1)
constexpr int YTH = YNW * YW;
constexpr int YST = 1024;  // tiles of 64 particles per pass-1 round
constexpr int YT = 4;      // tiles in flight per wavefront
constexpr int Y2T = 256;   // pass-2 workgroup
constexpr int Y2TAIL = 8192;  // top slots whose holes pass 2 marks in LDS

__device__ inline int lanes_below(unsigned long long m) {
  const int lane = threadIdx.x & 63;
  return __popcll(m & ((1ULL << lane) - 1ULL));
}

// Row change of particle p of row j, whose x and y the caller loaded: -1,
// 0, +1, or 2 (a jump the reference does not index).  FP64: jpos =
// int(y/delx) (boundary_periodic.f90:147), with particle_x applied first
// when fused: xmode 1 the periodic round-down wrap, 2 the walls' mirrors.
// Taking the loaded values lets a caller issue several particles' loads
// before the first store of a wrapped x.
__device__ inline int row_change(const Rec64 &R, size_t p, double x, double y, int j, int xmode,
                                 const BP &P, int &bad) {
  if (xmode == 2) {
    const int inew = wall_x_at(R, p, x, P);
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  } else if (xmode == 1) {
    double xv = x;
    wrap_x(xv, P);
    if (xv != x) R.x[p] = xv;
    const int inew = (int)(xv / P.delx);
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  }
  const int jpos = (int)(y / P.delx);
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

// Wavefront-inclusive prefix sum.
__device__ inline int wave_incl(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// pass 1: the tiles of a round are classified by all wavefronts (tile t by
// wavefront t % YNW), scanned, and their movers placed in index order.
template <class R>
__global__ __launch_bounds__(YTH) void k_particle_y1(BP P, int xmode,
                                                     const int *__restrict__ np2, R B,
                                                     const int64_t *__restrict__ id,
                                                     int *__restrict__ hole,
                                                     int *__restrict__ nhole,
                                                     int *__restrict__ nup,
                                                     int *__restrict__ ndn, R mu,
                                                     int64_t *mid_u, R md, int64_t *mid_d,
                                                     int *derr) {
  __shared__ unsigned long long smu[YST], smd[YST];   // per tile: lanes moving up / down
  __shared__ int scu[YST], scd[YST];                   // movers up / down before the tile
  __shared__ int wtu[YNW], wtd[YNW];
  const int seg = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int jl = seg % P.ny;
  const int j = P.nys + jl;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  const size_t mb = (size_t)seg * P.mcap;
  int cu = 0, cd = 0;                                  // movers of earlier rounds
  int bad = 0;
  for (int k0 = 0; k0 < n; k0 += YST * YW) {
    const int ntile = min(YST, (n - k0 + YW - 1) / YW);
    // classify (the periodic x wrap or the walls' mirrors are applied here)
    for (int t0 = w; t0 < ntile; t0 += YNW * YT) {
      // every tile's x and y in flight before the first wrapped x is stored
      double xv[YT], yv[YT];
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        const int k = k0 + (t0 + b * YNW) * YW + lane;
        xv[b] = yv[b] = 0.0;
        if (t0 + b * YNW < ntile && k < n) {
          xv[b] = B.x[base + k];
          yv[b] = B.y[base + k];
        }
      }
      int dir[YT];
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        const int k = k0 + (t0 + b * YNW) * YW + lane;
        dir[b] = t0 + b * YNW < ntile && k < n
                     ? row_change(B, base + k, xv[b], yv[b], j, xmode, P, bad)
                     : 0;
      }
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        const int t = t0 + b * YNW;
        if (dir[b] == 2) bad |= DERR_JUMP;
        const unsigned long long u = __ballot(dir[b] == 1), d = __ballot(dir[b] == -1);
        if (t < ntile && lane == 0) {
          smu[t] = u;
          smd[t] = d;
        }
      }
    }
    __syncthreads();
    // exclusive prefix of the tiles' counts: thread th owns tiles 2 th, 2 th + 1
    int a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    const int t2 = 2 * threadIdx.x;
    if (t2 < ntile) {
      a0 = __popcll(smu[t2]);
      b0 = __popcll(smd[t2]);
    }
    if (t2 + 1 < ntile) {
      a1 = __popcll(smu[t2 + 1]);
      b1 = __popcll(smd[t2 + 1]);
    }
    const int iu = wave_incl(a0 + a1), id_ = wave_incl(b0 + b1);
    if (lane == YW - 1) {
      wtu[w] = iu;
      wtd[w] = id_;
    }
    __syncthreads();
    int pu = 0, pd = 0, tu = 0, td = 0;
#pragma unroll
    for (int v = 0; v < YNW; ++v) {
      pu += v < w ? wtu[v] : 0;
      pd += v < w ? wtd[v] : 0;
      tu += wtu[v];
      td += wtd[v];
    }
    if (t2 < YST) {
      const int eu = cu + pu + iu - (a0 + a1), ed = cd + pd + id_ - (b0 + b1);
      scu[t2] = eu;
      scd[t2] = ed;
      scu[t2 + 1] = eu + a0;
      scd[t2 + 1] = ed + b0;
    }
    __syncthreads();
    // place each tile's movers
    for (int t = w; t < ntile; t += YNW) {
      const unsigned long long mu_ = smu[t], md_ = smd[t];
      if ((mu_ | md_) == 0ULL) continue;
      const bool up = (mu_ >> lane) & 1ULL, dn = (md_ >> lane) & 1ULL;
      if (up || dn) {
        const int k = k0 + t * YW + lane;
        // holes in ascending index order: flag(:,j) of :163-164
        const int hpos = scu[t] + scd[t] + lanes_below(mu_ | md_);
        if (hpos < 2 * P.mcap) hole[2 * mb + hpos] = k;
        else bad |= DERR_MOVER_OVER;
        const int pos = up ? scu[t] + lanes_below(mu_) : scd[t] + lanes_below(md_);
        if (pos < P.mcap) {
          const size_t q = mb + pos;
          (up ? mu : md).store(q, arriving(B, base + k, j, up ? 1 : -1, P));
          (up ? mid_u : mid_d)[q] = id[base + k];
        } else {
          bad |= DERR_MOVER_OVER;
        }
      }
    }
    cu += tu;
    cd += td;
    __syncthreads();                                   // the round's LDS is reused
  }
  if (threadIdx.x == 0) {
    nhole[seg] = min(cu + cd, 2 * P.mcap);
    nup[seg] = min(cu, P.mcap);
    ndn[seg] = min(cd, P.mcap);
  }
  if (bad) atomicOr(derr, bad);
}

// pass 2: one workgroup per segment; the arrivals are independent, and the
// first wavefront alone compacts the row when fewer arrive than left.
template <class R>
__global__ __launch_bounds__(Y2T) void k_particle_y2(BP P, int *__restrict__ np2, R B,
                                                    int64_t *__restrict__ id,
                                                    const int *__restrict__ hole,
                                                    const int *__restrict__ nhole,
                                                    const int *__restrict__ nup,
                                                    const int *__restrict__ ndn, R mu,
                                                    const int64_t *mid_u, R md,
                                                    const int64_t *mid_d, int *derr) {
  __shared__ unsigned char tail[Y2TAIL];           // holes among the top slots
  const int seg = blockIdx.x;
  const int lane = threadIdx.x;                    // the compaction's lanes: wavefront 0
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
  for (int a = threadIdx.x; a < A; a += Y2T) {
    const int dst = a < H ? hl[a] : N + (a - H);
    const bool f = a < nF;
    const size_t q = (size_t)(f ? Fseg : Sseg) * P.mcap + (f ? a : a - nF);
    B.store(base + dst, (f ? FB : SB).load(q));
    id[base + dst] = (f ? Fid : Sid)[q];
  }
  // With fewer arrivals than holes the top m = H - A slots empty: their
  // non-holes fill the holes the arrivals leave.  Sources are the non-holes
  // of [N - m, N) and destinations holes the arrivals do not take, so this
  // runs beside the arrivals' stores.  The holes among the top m slots are
  // the last entries of the ascending list (at most m of them, all in
  // hl[A..H)): the workgroup marks them in LDS first.
  const int m = H - A;
  const bool marked = A < H && m <= Y2TAIL;
  if (marked) {
    for (int i = threadIdx.x; i < m; i += Y2T) tail[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += Y2T) {
      const int v = hl[H - 1 - i];
      if (v >= N - m) tail[v - (N - m)] = 1;
    }
    __syncthreads();
  }
  if (A < H && threadIdx.x < YW) {
    int done = 0;    // non-hole tail slots already placed
    for (int q0 = 0; q0 < m; q0 += YW) {
      const int qi = q0 + lane;          // slot N-1-qi, walking down from the top
      bool nonhole = false;
      int src = 0;
      if (qi < m) {
        src = N - 1 - qi;
        if (marked) {
          nonhole = tail[m - 1 - qi] == 0;
        } else {
          // src is a hole iff it is in the sorted list hl[A..H)
          int lo = A, hi = H;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (hl[mid] < src) lo = mid + 1;
            else hi = mid;
          }
          nonhole = !(lo < H && hl[lo] == src);
        }
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
  __syncthreads();                                 // every lane has read np2[seg]
  if (threadIdx.x == 0) np2[seg] = newn;
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
    hipLaunchKernelGGL(k_particle_y1<Rec64>, dim3(nseg), dim3(YTH), 0, ctx->stream, P, xmode,
                       ctx->np2, B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, ctx->derr);
    hipLaunchKernelGGL(k_particle_y2<Rec64>, dim3(nseg), dim3(Y2T), 0, ctx->stream, P, ctx->np2,
                       B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  return WM_OK;
}
