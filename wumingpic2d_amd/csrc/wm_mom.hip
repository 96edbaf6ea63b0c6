// wm_mom.hip -- the moment diagnostics of an output step on the device:
// mom_calc__accl's velocity update over half a step (mom_calc.f90:48-164),
// mom_calc__nvt's cloud-in-cell moments N, V, T (:167-249), fused per tile of
// cells, and boundary_periodic__mom's periodic fold
// (boundary_periodic.f90:571-636).
//
// One workgroup takes MC cells x MR rows of one species.  Eight lanes share a
// cell and take its particles in rounds of eight, one per lane: each lane
// stages its particle's weights at the 3 x 3 nodes around the cell (x - 1/2
// truncates to the cell or the one before it) and its 7 moment values in
// LDS, then lane m adds moment m of the round's particles, in lane order, to
// its 9 node sums.  The cell sums go into the tile's node sums in LDS in
// ascending cell order.  The tiles overlap by two node columns and two node
// rows; a gather adds the (at most four) tiles of a node in a fixed order.
// Every sum has one order: the result does not depend on scheduling.
#include "wm_internal.h"

namespace {

constexpr int MC = 32;           // cells per workgroup along x
constexpr int MR = 4;            // rows per workgroup
constexpr int MG = 8;            // lanes per cell
constexpr int MW = MC + 2;       // tile columns: cells / nodes i0-1 .. i0+MC
constexpr int MH = MR + 2;       // tile rows: j0-1 .. j0+MR
constexpr int NMOM = 7;          // N, Vx, Vy, Vz, Txx, Tyy, Tzz
constexpr int NST = 6 + NMOM;    // staged per particle: 3 + 3 weights, 7 moment values
constexpr int TSZ = NMOM * MH * MW;
static_assert(9 * NMOM <= MG * NST, "a group's slots hold its cell's sums");

// The first LDS slot of group g.
__device__ inline double *grp_w(double *St, int g) { return St + g * MG * NST; }

struct MP {
  int nx, ny, nxgs, nys, nsp, nchunk, nrb, cap;
  double d_delx, c, delt;        // delt: the moment step, half the run's (mom_calc.f90:34)
  double q[WM_MAXSP], r[WM_MAXSP];
};

struct Pt {
  double x, y, ux, uy, uz;
};

// FP64 form: the records of cell i (local index) of a row, in cumcnt order.
struct Src64 {
  const double *x, *y, *ux, *uy, *uz;
  const int *cumcnt;
  __device__ void range(const MP &P, int s, int jl, int i, size_t &base, int &kb, int &ke) const {
    const size_t seg = (size_t)s * P.ny + jl;
    const int *cc = cumcnt + seg * (size_t)(P.nx + 1);
    kb = max(0, cc[i]);
    ke = min(P.cap, cc[i + 1]);
    base = seg * (size_t)P.cap;
  }
  __device__ Pt load(size_t p, int, int) const { return Pt{x[p], y[p], ux[p], uy[p], uz[p]}; }
};

// FP32 form: the cell's bin.  The position is the cell's index plus the
// offset, in FP64, as the layout conversion to the reference arrays forms it.
struct Src32 {
  const float4 *q;
  const float *uz;
  const int *cnt;
  const int *bstart;
  __device__ void range(const MP &P, int s, int jl, int i, size_t &base, int &kb, int &ke) const {
    const size_t seg = (size_t)s * P.ny + jl;
    kb = 0;
    ke = cnt[seg * P.nx + i];
    base = seg * (size_t)P.cap + bstart[seg * (P.nx + 1) + i];
  }
  __device__ Pt load(size_t p, int iabs, int jabs) const {
    const float4 v = q[p];
    return Pt{(double)iabs + (double)v.x, (double)jabs + (double)v.y, (double)v.z, (double)v.w,
              (double)uz[p]};
  }
};

__device__ inline void tsc(double dh, double &sm, double &s0, double &sp) {
  // mom_calc.f90:97-105
  sm = 0.5 * (0.5 - dh) * (0.5 - dh);
  s0 = 0.75 - dh * dh;
  sp = 0.5 * (0.5 + dh) * (0.5 + dh);
}

// mom_calc.f90:107-155: the Boris update of u over the moment step from the
// cell-centred fields F (rows of cells j-1..j+1 from Frow, columns of cells
// i-1..i+1 from col); the position is not advanced.
__device__ inline void half_push(const double *F, int Frow, int col, const Pt &v, int iabs,
                                 int jabs, const MP &P, double fac1, double txxx, double fac2,
                                 double &ux, double &uy, double &uz) {
  double sx0, sx1, sx2, sy0, sy1, sy2;
  tsc(v.x * P.d_delx - 0.5 - iabs, sx0, sx1, sx2);
  tsc(v.y * P.d_delx - 0.5 - jabs, sy0, sy1, sy2);
  double f[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double *T = F + (k * MH + Frow) * MW + col;
    f[k] = (T[0] * sx0 + T[1] * sx1 + T[2] * sx2) * sy0 +
           (T[MW] * sx0 + T[MW + 1] * sx1 + T[MW + 2] * sx2) * sy1 +
           (T[2 * MW] * sx0 + T[2 * MW + 1] * sx1 + T[2 * MW + 2] * sx2) * sy2;
  }
  const double bpx = f[0], bpy = f[1], bpz = f[2], epx = f[3], epy = f[4], epz = f[5];
  double uvm1 = v.ux + fac1 * epx;
  double uvm2 = v.uy + fac1 * epy;
  double uvm3 = v.uz + fac1 * epz;
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
  ux = uvm1 + fac1 * epx;
  uy = uvm2 + fac1 * epy;
  uz = uvm3 + fac1 * epz;
}

// Workgroup (chunk, row block, species).  ACCL: update u over the moment step
// first.  GP: write the updated records to the gp bank instead of summing
// moments (mom_calc__accl on its own).
template <class S, bool ACCL, bool GP>
__global__ __launch_bounds__(256) void k_mom(MP P, FieldGeom fg, const double *__restrict__ uf,
                                             S src, double *__restrict__ part, Rec64 gp,
                                             int *__restrict__ derr) {
  __shared__ double F[ACCL ? 6 * MH * MW : 1];
  __shared__ double Nt[GP ? 1 : TSZ];
  // per lane, one particle of a round: x weights, y weights, the 7 moment
  // values; after a row, each group's slots hold its cell's 63 sums
  __shared__ double St[GP ? 1 : MC * MG * NST];
  const int chunk = blockIdx.x, rb = blockIdx.y, s = blockIdx.z;
  const int c0 = chunk * MC, r0 = rb * MR;       // first cell, first row (local indices)
  if (ACCL) {
    // cell-centred fields, mom_calc.f90:68-81, of cells c0-1 .. c0+MC and
    // rows r0-1 .. r0+MR; cell (a, b) is at (a+2, b+2) of a ghost-inclusive plane
    for (int t = threadIdx.x; t < 6 * MH * MW; t += blockDim.x) {
      const int col = t % MW, row = (t / MW) % MH, k = t / (MW * MH);
      const int a = c0 - 1 + col, b = r0 - 1 + row;
      double v = 0.0;
      if (a <= P.nx && b <= P.ny) {
        const size_t gx = fg.gx, o = (size_t)(a + 2) + gx * (size_t)(b + 2);
        const double *U = uf + (size_t)k * fg.ps;
        switch (k) {
          case 0: v = 0.5 * (U[o] + U[o + gx]); break;
          case 1: v = 0.5 * (U[o] + U[o + 1]); break;
          case 2: v = 0.25 * (U[o] + U[o + 1] + U[o + gx] + U[o + gx + 1]); break;
          case 3: v = 0.5 * (U[o] + U[o + 1]); break;
          case 4: v = 0.5 * (U[o] + U[o + gx]); break;
          default: v = U[o]; break;
        }
      }
      F[t] = v;
    }
  }
  if (!GP)
    for (int t = threadIdx.x; t < TSZ; t += blockDim.x) Nt[t] = 0.0;
  __syncthreads();
  const int g = threadIdx.x / MG, sub = threadIdx.x % MG;
  const int i = c0 + g;
  const double q = P.q[s], rr = P.r[s];
  const double fac1 = q / rr * 0.5 * P.delt;     // mom_calc.f90:90-92
  const double txxx = fac1 * fac1;
  const double fac2 = q * P.delt / rr;
  const double ic2 = 1.0 / (P.c * P.c);
  bool bad = false;
  double *my = St + (g * MG + sub) * NST;        // this lane's staging slot
  const double *grp = St + g * MG * NST;         // the group's eight slots
  for (int rl = 0; rl < MR; ++rl) {
    const int jl = r0 + rl;
    if (jl >= P.ny) break;                       // the same for the whole workgroup
    const int iabs = P.nxgs + i, jabs = P.nys + jl;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; // moment `sub` at the 3 x 3 nodes
    if (i < P.nx) {
      size_t base;
      int kb, ke;
      src.range(P, s, jl, i, base, kb, ke);
      // rounds of eight particles, one per lane; the next round's particle
      // is loaded while this one is computed
      int k = kb + sub;
      Pt nxt{0, 0, 0, 0, 0};
      if (k < ke) nxt = src.load(base + (size_t)k, iabs, jabs);
      for (int k0 = kb; k0 < ke; k0 += MG, k += MG) {
        const Pt v = nxt;
        const bool live = k < ke;
        if (k + MG < ke) nxt = src.load(base + (size_t)(k + MG), iabs, jabs);
        double ux = v.ux, uy = v.uy, uz = v.uz;
        if (ACCL && live) half_push(F, rl, g, v, iabs, jabs, P, fac1, txxx, fac2, ux, uy, uz);
        if (GP) {
          if (live) {
            const size_t p = base + (size_t)k;
            gp.x[p] = v.x;
            gp.y[p] = v.y;
            gp.ux[p] = ux;
            gp.uy[p] = uy;
            gp.uz[p] = uz;
          }
          continue;
        }
        // mom_calc.f90:190-201: ih = int(x - 1/2) (truncation), the cell or
        // the one before it for a particle of this cell
        double w[NST];
#pragma unroll
        for (int e = 0; e < NST; ++e) w[e] = 0.0;
        const double xs = v.x * P.d_delx - 0.5, ys = v.y * P.d_delx - 0.5;
        const int ih = (int)xs, jh = (int)ys;
        const int di = ih - iabs, dj = jh - jabs;
        const bool inside = di >= -1 && di <= 0 && dj >= -1 && dj <= 0;
        bad |= live && !inside;
        if (live && inside) {
          const double dx = xs - ih, dxm = 1.0 - dx, dy = ys - jh, dym = 1.0 - dy;
          w[0] = di < 0 ? dxm : 0.0;             // x weights of nodes i-1, i, i+1
          w[1] = di < 0 ? dx : dxm;
          w[2] = di < 0 ? 0.0 : dx;
          w[3] = dj < 0 ? dym : 0.0;             // y weights of nodes j-1, j, j+1
          w[4] = dj < 0 ? dy : dym;
          w[5] = dj < 0 ? 0.0 : dy;
          const double ig = 1.0 / sqrt(1.0 + (ux * ux + uy * uy + uz * uz) * ic2);
          w[6] = 1.0;                            // N, V, T: mom_calc.f90:203-243
          w[7] = ux * ig;
          w[8] = uy * ig;
          w[9] = uz * ig;
          w[10] = ux * ux * ig;
          w[11] = uy * uy * ig;
          w[12] = uz * uz * ig;
        }
#pragma unroll
        for (int e = 0; e < NST; ++e) my[e] = w[e];
        // the group's eight lanes are in one wavefront, whose LDS accesses
        // complete in order: the slots are written before any lane reads them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (sub < NMOM) {
          // lane m adds moment m of the round's particles, in lane order:
          // ((v_m x-weight) y-weight), the reference's rounding per term
#pragma unroll
          for (int j = 0; j < MG; ++j) {
            const double *o = grp + j * NST;
            const double vm = o[6 + sub];
            const double t0 = vm * o[0], t1 = vm * o[1], t2 = vm * o[2];
            const double y0 = o[3], y1 = o[4], y2 = o[5];
            acc[0] += t0 * y0; acc[1] += t1 * y0; acc[2] += t2 * y0;
            acc[3] += t0 * y1; acc[4] += t1 * y1; acc[5] += t2 * y1;
            acc[6] += t0 * y2; acc[7] += t1 * y2; acc[8] += t2 * y2;
          }
        }
        // every read of this round is issued before the next round writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (GP) continue;
    // the cell's sums into the group's own slots: [(b*3 + a)*NMOM + m]
    if (sub < NMOM) {
#pragma unroll
      for (int e = 0; e < 9; ++e) grp_w(St, g)[e * NMOM + sub] = acc[e];
    }
    __syncthreads();
    // node (tile column c, node row j-1+b) takes cells c-2, c-1, c in order
    for (int t = threadIdx.x; t < NMOM * 3 * MW; t += blockDim.x) {
      const int c = t % MW, b = (t / MW) % 3, m = t / (3 * MW);
      double *d = Nt + (m * MH + rl + b) * MW + c;
      double sum = *d;
      for (int gg = max(0, c - 2); gg <= min(MC - 1, c); ++gg)
        sum += St[gg * MG * NST + (b * 3 + (c - gg)) * NMOM + m];
      *d = sum;
    }
    __syncthreads();
  }
  if (!GP) {
    double *T = part + (((size_t)s * P.nrb + rb) * P.nchunk + chunk) * TSZ;
    for (int t = threadIdx.x; t < TSZ; t += blockDim.x) T[t] = Nt[t];
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

// mom (7, nxgs-1:nxge+1, nys-1:nye+1, nsp) from the tiles: node (a, b)
// (offsets from nxgs-1, nys-1) is column a - ch*MC of chunk ch and row
// b - rb*MR of row block rb; the tiles add in ascending (rb, ch).
__global__ void k_mom_gather(MP P, const double *__restrict__ part, double *__restrict__ mom) {
  const int NXm = P.nx + 2, NYm = P.ny + 2;
  const size_t n = (size_t)NMOM * NXm * NYm * P.nsp;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int m = (int)(t % NMOM);
  size_t rest = t / NMOM;
  const int a = (int)(rest % NXm);
  rest /= NXm;
  const int b = (int)(rest % NYm);
  const int s = (int)(rest / NYm);
  double sum = 0.0;
  for (int rb = max(0, b / MR - 1); rb <= min(P.nrb - 1, b / MR); ++rb) {
    const int tr = b - rb * MR;
    if (tr >= MH) continue;
    for (int ch = max(0, a / MC - 1); ch <= min(P.nchunk - 1, a / MC); ++ch) {
      const int tc = a - ch * MC;
      if (tc >= MW) continue;
      sum += part[(((size_t)s * P.nrb + rb) * P.nchunk + ch) * TSZ + ((size_t)m * MH + tr) * MW + tc];
    }
  }
  mom[t] = sum;
}

// boundary_periodic.f90:579-584: the x ghosts into the opposite edge, every
// row; the wall families fold each into its own edge (boundary_shock.f90:
// 636-641, boundary_reconnection.f90:590-595).
__global__ void k_mom_fold_x(int nx, int ny, int nsp, double *mom, int wall) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NMOM * (ny + 2) * nsp) return;
  const int m = t % NMOM, row = t / NMOM;        // row = b + (ny+2) * s
  double *r = mom + m + (size_t)NMOM * (nx + 2) * row;
  const double lo = r[0], hi = r[NMOM * (nx + 1)];
  r[NMOM * 1] += wall ? lo : hi;
  r[NMOM * nx] += wall ? hi : lo;
}

// boundary_periodic.f90:586-633 on one rank: row nys-1 into row nye, then
// row nye+1 into row nys, every column.
__global__ void k_mom_fold_y(int nx, int ny, int nsp, double *mom) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NMOM * (nx + 2) * nsp) return;
  const int e = t % (NMOM * (nx + 2)), s = t / (NMOM * (nx + 2));
  const size_t rs = (size_t)NMOM * (nx + 2);
  double *c = mom + e + rs * (ny + 2) * s;
  c[rs * ny] += c[0];
  c[rs * 1] += c[rs * (ny + 1)];
}

MP make_mp(const wm_ctx *ctx) {
  const wm_grid &g = ctx->g;
  MP P;
  P.nx = g.nx;
  P.ny = g.ny;
  P.nxgs = g.nxgs;
  P.nys = g.nys;
  P.nsp = g.nsp;
  P.nchunk = wm_cdiv(g.nx, MC);
  P.nrb = wm_cdiv(g.ny, MR);
  P.cap = ctx->cap;
  P.d_delx = 1. / g.delx;
  P.c = g.c;
  P.delt = g.delt * 0.5;
  for (int s = 0; s < WM_MAXSP; ++s) {
    P.q[s] = s < g.nsp ? g.q[s] : 0.0;
    P.r[s] = s < g.nsp ? g.r[s] : 1.0;
  }
  return P;
}

template <class S, bool ACCL, bool GP>
int launch_mom(wm_ctx *ctx, const MP &P, const S &src, Rec64 gp) {
  hipLaunchKernelGGL((k_mom<S, ACCL, GP>), dim3(P.nchunk, P.nrb, P.nsp), dim3(MC * MG), 0,
                     ctx->stream, P, ctx->fg, ctx->uf, src, ctx->mom_part, gp, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

}  // namespace

size_t wm_mom_elems(const wm_ctx *ctx) {
  return (size_t)NMOM * (ctx->g.nx + 2) * (ctx->g.ny + 2) * ctx->g.nsp;
}

int wm_mom_buffers(wm_ctx *ctx) {
  const MP P = make_mp(ctx);
  const size_t np = (size_t)P.nsp * P.nrb * P.nchunk * TSZ;
  if (!ctx->mom) WM_HIP(hipMalloc((void **)&ctx->mom, sizeof(double) * wm_mom_elems(ctx)));
  if (!ctx->mom_part) WM_HIP(hipMalloc((void **)&ctx->mom_part, sizeof(double) * np));
  return WM_OK;
}

// Into ctx->mom, allocated here on first use (read the pointer only after).
int wm_k_moments(wm_ctx *ctx, bool accl) {
  WM_TRY(wm_mom_buffers(ctx));
  const MP P = make_mp(ctx);
  if (ctx->form == WM_FORM_FP32) {
    if (!accl) return WM_ERR_ARG;
    const Src32 src{ctx->bins.q, ctx->bins.uz, ctx->bin_cnt, ctx->bstart};
    WM_TRY((launch_mom<Src32, true, false>(ctx, P, src, Rec64{})));
  } else {
    const PBank &A = ctx->bank[0];
    const Src64 src{A.x, A.y, A.ux, A.uy, A.uz, ctx->cumcnt};
    if (accl)
      WM_TRY((launch_mom<Src64, true, false>(ctx, P, src, Rec64{})));
    else
      WM_TRY((launch_mom<Src64, false, false>(ctx, P, src, Rec64{})));
  }
  const size_t n = wm_mom_elems(ctx);
  hipLaunchKernelGGL(k_mom_gather, dim3(wm_cdiv((long long)n, 256)), dim3(256), 0, ctx->stream, P,
                     ctx->mom_part, ctx->mom);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_accl(wm_ctx *ctx, int src_bank, int dst_bank) {
  const MP P = make_mp(ctx);
  const PBank &A = ctx->bank[src_bank];
  const Src64 src{A.x, A.y, A.ux, A.uy, A.uz, ctx->cumcnt};
  return launch_mom<Src64, true, true>(ctx, P, src, ctx->bank[dst_bank]);
}

int wm_k_bc_mom(wm_ctx *ctx, double *mom) {
  const wm_grid &g = ctx->g;
  const int nxr = NMOM * (g.ny + 2) * g.nsp, nyr = NMOM * (g.nx + 2) * g.nsp;
  hipLaunchKernelGGL(k_mom_fold_x, dim3(wm_cdiv(nxr, 256)), dim3(256), 0, ctx->stream, g.nx, g.ny,
                     g.nsp, mom, ctx->bcf != WM_BC_PERIODIC);
  WM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mom_fold_y, dim3(wm_cdiv(nyr, 256)), dim3(256), 0, ctx->stream, g.nx, g.ny,
                     g.nsp, mom);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
