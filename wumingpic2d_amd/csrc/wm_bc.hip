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
#include "wm_bins.h"

namespace {

struct BP {
  int nx, ny, nxgs, nxge, nygs, nyge, nys, nsp, cap, mcap;
  double delx, spanx, spany;
  int unit;            // delx == 1 (the reference's apps): v / delx is v itself
  int xs, xe;          // active columns nxs, nxe (the wall families)
  double wl, wr;       // mirror points 2 (nxs+1) delx and 2 (nxe-1) delx
};

// The outputs pass 1 gives the fused pass 2 + sort (all null otherwise).
struct YFuse {
  unsigned long long *tmask;    // per tile of 64 slots: the lanes that leave
  int *tpre;                    // per tile: the row's holes before it
  int ntiles;                   // tiles per segment
  int *hist;                    // [seg][nx]: cells after pass 2 (added to, zero before)
};

// int(v / delx) of the reference: with delx == 1 the quotient is v exactly
// (a uniform branch on the kernel argument skips the division).
__device__ inline int cell_div(double v, const BP &P) {
  if (P.unit) return (int)v;
  return (int)(v / P.delx);
}

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
  const int ipos = cell_div(x, P);
  if (ipos < P.nxgs) x = add_rd(x, P.spanx);
  else if (ipos >= P.nxge + 1) x = add_rd(x, -P.spanx);
  return ipos;
}

// The wall families' particle_x (boundary_shock.f90:77-89,
// boundary_reconnection.f90:76-88): a particle in a cell below nxs+1 is
// mirrored about (nxs+1) delx, one in a cell at or above nxe-1 about
// (nxe-1) delx, and its momentum reversed.  Returns the new cell.
__device__ inline int wall_x_at(const Rec64 &R, size_t p, double x, const BP &P, double &xf) {
  const int ipos = cell_div(x, P);
  double w;
  xf = x;
  if (ipos < P.xs + 1) w = P.wl;
  else if (ipos >= P.xe - 1) w = P.wr;
  else return ipos;
  const double xn = w - x;
  xf = xn;
  R.x[p] = xn;
  R.ux[p] = -R.ux[p];
  R.uy[p] = -R.uy[p];
  R.uz[p] = -R.uz[p];
  return cell_div(xn, P);
}
__device__ inline int wall_x(const Rec64 &R, size_t p, const BP &P) {
  double xf;
  return wall_x_at(R, p, R.x[p], P, xf);
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
    const int ipos = cell_div(x, P);
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
    const int inew = cell_div(xn, P);
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
    const int inew = cell_div(v, P);
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
                                 const BP &P, int &bad, double &xf) {
  xf = x;                                          // the particle's x once particle_x is done
  if (xmode == 2) {
    const int inew = wall_x_at(R, p, x, P, xf);
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  } else if (xmode == 1) {
    double xv = x;
    int inew = wrap_x(xv, P);                      // int(x / delx), unless it wrapped
    if (xv != x) {
      R.x[p] = xv;
      inew = cell_div(xv, P);
    }
    xf = xv;
    if (inew < P.nxgs || inew > P.nxge) bad |= DERR_RANGE;
  }
  const int jpos = cell_div(y, P);
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
                                                     YFuse F, int *derr) {
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
    // the next YT tiles' x and y are in flight while these are classified;
    // indices are clamped into the segment (lanes past its end read its last
    // particle and are not classified), so the loads are unconditional
    double xv[YT], yv[YT];
    auto fetch = [&](int t0, double *xs, double *ys) {
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        const int k = min(k0 + (t0 + b * YNW) * YW + lane, n - 1);
        xs[b] = B.x[base + k];
        ys[b] = B.y[base + k];
      }
    };
    fetch(w, xv, yv);
    for (int t0 = w; t0 < ntile; t0 += YNW * YT) {
      double xn[YT], yn[YT];
      fetch(t0 + YNW * YT, xn, yn);
      int dir[YT];
      double xf[YT];
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        const int k = k0 + (t0 + b * YNW) * YW + lane;
        xf[b] = 0.0;
        dir[b] = t0 + b * YNW < ntile && k < n
                     ? row_change(B, base + k, xv[b], yv[b], j, xmode, P, bad, xf[b])
                     : 0;
      }
      if (F.hist) {
        // the histogram of the cells each row holds after pass 2 (the sort's
        // count, int(x) as sort.f90:60): a staying particle's in this row's,
        // one atomic per distinct cell of the wavefront; a leaver's in the
        // receiving row's
#pragma unroll
        for (int b = 0; b < YT; ++b) {
          const int k = k0 + (t0 + b * YNW) * YW + lane;
          const bool in = t0 + b * YNW < ntile && k < n;
          const int key = (int)xf[b] - P.nxgs;
          const bool ok = in && key >= 0 && key < P.nx;
          const bool stay = ok && dir[b] == 0;
          const wm_bins::WaveMatch mm = wm_bins::wave_match(key, stay);
          if (stay && mm.leader) atomicAdd(F.hist + (size_t)seg * P.nx + key, mm.count);
          if (ok && (dir[b] == 1 || dir[b] == -1)) {
            const int s = seg / P.ny, dj = wm_wrap(jl + dir[b], P.ny);
            atomicAdd(F.hist + ((size_t)s * P.ny + dj) * P.nx + key, 1);
          }
        }
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
#pragma unroll
      for (int b = 0; b < YT; ++b) {
        xv[b] = xn[b];
        yv[b] = yn[b];
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
      if (F.tmask) {                                   // each tile's leavers and the holes before it
        const size_t tb = (size_t)seg * F.ntiles + k0 / YW;
        if (t2 < ntile) {
          F.tmask[tb + t2] = smu[t2] | smd[t2];
          F.tpre[tb + t2] = eu + ed;
        }
        if (t2 + 1 < ntile) {
          F.tmask[tb + t2 + 1] = smu[t2 + 1] | smd[t2 + 1];
          F.tpre[tb + t2 + 1] = eu + a0 + ed + b0;
        }
      }
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

// pass 2 fused with the sort (the resident step; sort.f90:36-82 after
// bp.f90:143-236).  The row pass 2 would leave -- hole h filled by arrival h
// while there are arrivals, surplus arrivals appended, or with fewer
// arrivals the top m = H - A slots' non-holes moved, from the top, into the
// holes the arrivals leave -- is never written: slot k of it is read where
// it comes from (the bank, an arrival list, or the slot the compaction
// moves), through pass 1's tile masks and hole prefixes, and sorted straight
// into the other bank.  The counts come from pass 1's histogram (cleared
// here for the next step); the scatter is k_sort's: rounds of FW tiles, one
// per wavefront, ranks among equal keys by a ballot loop, each wavefront's
// (key, count) list in LDS for the lanes of the higher tiles.
constexpr int FW = 8;      // wavefronts per segment
constexpr int FTH = FW * YW;

template <class R>
__global__ __launch_bounds__(FTH) void k_y2_sort(BP P, int *__restrict__ np2, R B,
                                                 const int64_t *__restrict__ id,
                                                 const int *__restrict__ hole,
                                                 const int *__restrict__ nhole,
                                                 const int *__restrict__ nup,
                                                 const int *__restrict__ ndn, R mu,
                                                 const int64_t *mid_u, R md, const int64_t *mid_d,
                                                 YFuse F, R dst, int64_t *__restrict__ idd,
                                                 int *__restrict__ cumcnt, int *derr) {
  extern __shared__ int cnt[];                        // nx running offsets
  __shared__ int lkey[FW][YW], lcnt[FW][YW], nl[FW], wsum[FW];
  __shared__ int csrc[Y2TAIL];                        // the compaction's sources, in hole order
  using V = typename R::Val;
  const int seg = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int s = seg / P.ny, jl = seg % P.ny;
  const int below = s * P.ny + (jl == 0 ? P.ny - 1 : jl - 1);
  const int above = s * P.ny + (jl == P.ny - 1 ? 0 : jl + 1);
  // arrival lists, in the order the self-SENDRECVs append them (:174-189)
  const bool first_row = (jl == 0);
  const R FB = first_row ? md : mu, SB = first_row ? mu : md;
  const int64_t *Fid = first_row ? mid_d : mid_u, *Sid = first_row ? mid_u : mid_d;
  const int Fseg = first_row ? above : below, Sseg = first_row ? below : above;
  const int nF = first_row ? ndn[above] : nup[below];
  const int nS = first_row ? nup[below] : ndn[above];
  const int A = nF + nS, H = nhole[seg], N = np2[seg];
  const int newn = N + A - H, m = H - A;
  const size_t base = (size_t)seg * P.cap;
  const int *hl = hole + 2 * (size_t)seg * P.mcap;
  const unsigned long long *tm = F.tmask + (size_t)seg * F.ntiles;
  const int *tp = F.tpre + (size_t)seg * F.ntiles;
  int *hist = F.hist + (size_t)seg * P.nx;
  if (newn > P.cap) {
    if (threadIdx.x == 0) atomicOr(derr, DERR_NP_OVER);
    return;
  }
  // is slot k (< N) a hole of pass 1, and its rank among the holes
  auto hole_rank = [&](int k, int &h) {
    const unsigned long long mk = tm[k >> 6];
    const int b = k & 63;
    h = tp[k >> 6] + __popcll(mk & ((1ULL << b) - 1ULL));
    return (bool)((mk >> b) & 1ULL);
  };
  // the compaction's sources: the non-holes of [newn, N) from the top, in
  // order, for the holes of rank A, A + 1, ... below newn
  const bool in_lds = m <= Y2TAIL;
  if (m > 0 && in_lds) {
    int run = 0;
    for (int j0 = 0; j0 < m; j0 += FTH) {
      const int j = j0 + (int)threadIdx.x;
      int h;
      const bool nh = j < m && !hole_rank(N - 1 - j, h);
      const unsigned long long bm = __ballot(nh);
      if (lane == 0) wsum[w] = __popcll(bm);
      __syncthreads();
      int pre = 0, tot = 0;
#pragma unroll
      for (int v = 0; v < FW; ++v) {
        pre += v < w ? wsum[v] : 0;
        tot += wsum[v];
      }
      if (nh) csrc[run + pre + lanes_below(bm)] = N - 1 - j;
      __syncthreads();
      run += tot;
    }
  }
  // the r-th source without the LDS list (more than Y2TAIL slots vacated):
  // the smallest j with r + 1 non-holes among slots N-1 .. N-1-j
  auto source_far = [&](int r) {
    int lo = r, hi = m - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1, x = N - 1 - mid;
      int a = 0, b = H;                              // holes >= x: H - (first hole >= x)
      while (a < b) {
        const int c = (a + b) >> 1;
        if (hl[c] < x) a = c + 1;
        else b = c;
      }
      if (mid + 1 - (H - a) >= r + 1) hi = mid;
      else lo = mid + 1;
    }
    return N - 1 - lo;
  };
  // where slot k of the row after pass 2 comes from: the bank (returns the
  // slot) or an arrival list (returns -1 - its record, F or S)
  auto source = [&](int k, bool &fl, size_t &q) {
    int h = 0;
    int a = -1;
    if (k >= N) {
      a = H + (k - N);
    } else if (hole_rank(k, h)) {
      if (h < A) a = h;
      else if (m > 0 && h - A < m) return in_lds ? csrc[h - A] : source_far(h - A);
      else return -2;                                // inconsistent lists (a mover overflow)
    } else {
      return k;
    }
    fl = a < nF;
    q = fl ? (size_t)Fseg * P.mcap + a : (size_t)Sseg * P.mcap + (a - nF);
    return -1;
  };
  auto fetch = [&](int k, V &v, int64_t &vid) {
    bool fl = false;
    size_t q = 0;
    int sl = source(k, fl, q);
    if (sl == -2 || sl >= N) {                         // (only after a reported overflow)
      atomicOr(derr, DERR_MOVER_OVER);
      sl = 0;
    }
    if (sl >= 0) {
      v = B.load(base + sl);
      vid = id[base + sl];
    } else {
      v = (fl ? FB : SB).load(q);
      vid = (fl ? Fid : Sid)[q];
    }
  };
  // exclusive prefix of the histogram -> cumcnt and the running offsets;
  // the histogram is left zero for the next step
  {
    const int per = (P.nx + FTH - 1) / FTH;
    const int lo = min(P.nx, (int)threadIdx.x * per), hi = min(P.nx, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += hist[i];
    int incl = sum;
#pragma unroll
    for (int off = 1; off < YW; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    if (lane == YW - 1) wsum[w] = incl;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < FW; ++v) {
      pre += v < w ? wsum[v] : 0;
      tot += wsum[v];
    }
    int run = pre + incl - sum;
    int *cc = cumcnt + (size_t)seg * (P.nx + 1);
    for (int i = lo; i < hi; ++i) {
      const int c = hist[i];
      hist[i] = 0;
      cc[i] = run;
      cnt[i] = run;
      run += c;
    }
    if (threadIdx.x == 0) {
      cc[P.nx] = tot;
      if (tot != newn) atomicOr(derr, DERR_RANGE);   // a particle outside the row's cells
    }
  }
  __syncthreads();
  // stable scatter (sort.f90:71-75) of slots 0 .. newn-1 of the row after
  // pass 2, rounds of FW tiles, the next round's tile loading meanwhile
  const int nt = (newn + YW - 1) / YW;
  V v{}, nv{};
  int64_t vid = 0, nvid = 0;
  int key = -1, nkey = -1;
  {
    const int k = w * YW + lane;
    if (k < newn) {
      fetch(k, v, vid);
      key = B.key(v, P.nxgs);
    }
  }
  for (int r0 = 0; r0 < nt; r0 += FW) {
    {
      const int k = (r0 + FW + w) * YW + lane;
      nkey = -1;
      if (k < newn) {
        fetch(k, nv, nvid);
        nkey = B.key(nv, P.nxgs);
      }
    }
    const bool act = key >= 0 && key < P.nx;
    const wm_bins::WaveMatch mt = wm_bins::wave_match(key, act);
    const unsigned long long lm = __ballot(act && mt.leader);
    if (act && mt.leader) {
      const int li = __popcll(lm & ((1ULL << lane) - 1ULL));
      lkey[w][li] = key;
      lcnt[w][li] = mt.count;
    }
    if (lane == 0) nl[w] = __popcll(lm);
    __syncthreads();
    if (act) {
      int prior = 0;                                   // equal keys in the round's lower tiles
      for (int u = 0; u < w; ++u) {
        const int nu = nl[u];
        for (int i = 0; i < nu; ++i) prior += lkey[u][i] == key ? lcnt[u][i] : 0;
      }
      const int dpos = cnt[key] + prior + mt.rank;
      if (dpos < newn) {
        const size_t d = base + dpos;
        dst.store(d, v);
        idd[d] = vid;
      } else {
        atomicOr(derr, DERR_RANGE);
      }
    }
    __syncthreads();                                   // every lane has read the offsets and lists
    if (act && mt.leader) atomicAdd(&cnt[key], mt.count);
    v = nv;
    vid = nvid;
    key = nkey;
  }
  __syncthreads();                                     // every lane has read np2[seg]
  if (threadIdx.x == 0) np2[seg] = newn;
}

BP make_bp(const wm_ctx *ctx) {
  const wm_grid &g = ctx->g;
  BP P;
  P.nx = g.nx; P.ny = g.ny; P.nxgs = g.nxgs; P.nxge = g.nxge; P.nygs = g.nygs;
  P.nyge = g.nyge; P.nys = g.nys; P.nsp = g.nsp; P.cap = ctx->cap; P.mcap = ctx->mcap;
  P.delx = g.delx;
  P.unit = g.delx == 1.0;
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

int wm_k_particle_y_sort(wm_ctx *ctx, int src, int dst, int xmode, hipEvent_t mid) {
  BP P = make_bp(ctx);
  const int nseg = ctx->g.nsp * ctx->g.ny;
  const size_t lds = sizeof(int) * (size_t)P.nx;
  if (ctx->form == WM_FORM_FP32 || lds > 65536) {
    wm_set_error(ctx, WM_ERR_ARG, "particle_y + sort: FP64 rows of at most 16384 cells");
    return WM_ERR_ARG;
  }
  int64_t *id = ctx->id[ctx->idcur];
  const YFuse F{ctx->ytmask, ctx->ytpre, ctx->ytiles, ctx->yhist};
  const Rec64 &B = ctx->bank[src];
  hipLaunchKernelGGL(k_particle_y1<Rec64>, dim3(nseg), dim3(YTH), 0, ctx->stream, P, xmode,
                     ctx->np2, B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                     ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, F, ctx->derr);
  if (mid) WM_HIP(hipEventRecord(mid, ctx->stream));
  hipLaunchKernelGGL(k_y2_sort<Rec64>, dim3(nseg), dim3(FTH), lds, ctx->stream, P, ctx->np2, B,
                     id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up, ctx->mvid_up,
                     ctx->mv_dn, ctx->mvid_dn, F, ctx->b64(dst), ctx->id[1 - ctx->idcur],
                     ctx->cumcnt, ctx->derr);
  WM_HIP(hipGetLastError());
  ctx->idcur = 1 - ctx->idcur;
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
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, YFuse{}, ctx->derr);
    hipLaunchKernelGGL(k_particle_y2<Rec64>, dim3(nseg), dim3(Y2T), 0, ctx->stream, P, ctx->np2,
                       B, id, ctx->hole, ctx->nhole, ctx->nup, ctx->ndn, ctx->mv_up,
                       ctx->mvid_up, ctx->mv_dn, ctx->mvid_dn, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  return WM_OK;
}
