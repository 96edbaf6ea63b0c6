// wm_sort.hip -- stable per-row bucket sort by x-cell (sort.f90:36-82).
//
// One wavefront per (species, row) segment, so the ordered walk through the
// segment needs no workgroup barriers: the cell histogram and the running
// write offsets live in LDS (nx counters), and the wavefront walks the
// segment in index order, 64 particles at a time, with the loads of four
// such tiles in flight together.  Inside a tile, lanes with equal keys are
// grouped by a ballot loop over the distinct keys (sorted input has one to
// three per tile), which gives each lane its rank among equal keys below it,
// so the destination of every particle is
//   start[key] + (equal keys already written) + rank,
// exactly the reference's stable scatter.  Keys are int(x) as in sort.f90:60
// (the reference assumes delx = 1).
#include <algorithm>
#include <utility>

#include "wm_internal.h"

namespace {

constexpr int SW = 64;     // one wavefront per segment
constexpr int BT = 4;      // tiles in flight

struct SP {
  int nx, nxgs, cap;
};

struct WaveMatch {
  int rank;     // equal keys in lower lanes
  int count;    // equal keys in the wavefront
  bool leader;  // lowest lane holding the key
};

__device__ inline WaveMatch wave_match(int key, bool active) {
  const int lane = threadIdx.x & 63;
  WaveMatch w{0, 0, false};
  unsigned long long rem = __ballot(active);
  while (rem) {
    const int l = __ffsll((long long)rem) - 1;
    const int k = __shfl(key, l);
    const unsigned long long m = __ballot(active && key == k);
    if (active && key == k) {
      w.rank = __popcll(m & ((1ULL << lane) - 1ULL));
      w.count = __popcll(m);
      w.leader = (lane == l);
    }
    rem &= ~m;
  }
  return w;
}

template <class R>
__global__ __launch_bounds__(SW) void k_sort(SP P, const int *__restrict__ np2, R src,
                                             const int64_t *__restrict__ ids, R dst,
                                             int64_t *__restrict__ idd,
                                             int *__restrict__ cumcnt, int *derr) {
  extern __shared__ int cnt[];             // nx counters, then running offsets
  using V = typename R::Val;
  const int seg = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  int bad = 0;
  for (int i = lane; i < P.nx; i += SW) cnt[i] = 0;
  __syncthreads();
  // count (sort.f90:59-62)
  for (int k0 = 0; k0 < n; k0 += SW * BT) {
    int key[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = k0 + b * SW + lane;
      key[b] = k < n ? src.key(src.load(base + k), P.nxgs) : -1;
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = k0 + b * SW + lane;
      bool act = k < n;
      if (act && (key[b] < 0 || key[b] >= P.nx)) {
        bad = 1;
        act = false;
      }
      const WaveMatch m = wave_match(key[b], act);
      if (act && m.leader) cnt[key[b]] += m.count;
    }
  }
  __syncthreads();
  // exclusive prefix -> cumcnt(nxs:nxe+1) (sort.f90:64-69); cnt becomes the
  // running write offset of each cell
  {
    const int per = (P.nx + SW - 1) / SW;
    const int lo = min(P.nx, lane * per), hi = min(P.nx, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    int incl = s;
    for (int off = 1; off < SW; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    int run = incl - s;
    int *cc = cumcnt + (size_t)seg * (P.nx + 1);
    for (int i = lo; i < hi; ++i) {
      const int c = cnt[i];
      cc[i] = run;
      cnt[i] = run;
      run += c;
    }
    if (lane == SW - 1) cc[P.nx] = incl;
  }
  __syncthreads();
  // stable scatter (sort.f90:71-75), in index order
  for (int k0 = 0; k0 < n; k0 += SW * BT) {
    int key[BT];
    V v[BT];
    int64_t vid[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = k0 + b * SW + lane;
      key[b] = -1;
      if (k < n) {
        const size_t s = base + k;
        v[b] = src.load(s);
        vid[b] = ids[s];
        key[b] = src.key(v[b], P.nxgs);
      }
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const bool act = key[b] >= 0 && key[b] < P.nx;
      const WaveMatch m = wave_match(key[b], act);
      if (act) {
        const size_t d = base + cnt[key[b]] + m.rank;
        dst.store(d, v[b]);
        idd[d] = vid[b];
      }
      // every lane has read its offset before the leaders advance them
      __builtin_amdgcn_wave_barrier();
      if (act && m.leader) cnt[key[b]] += m.count;
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

// ------------------------------------------------------------------------
// FP32 form: row migration and cell sort as one stable counting scatter.
//
// The push counts, per source segment, where its particles go: bcnt[src
// seg][dy + 1][new cell] (dy = row change).  Destination row jd takes, cell
// by cell, the arrivals from row jd - 1 (moving up), then the particles that
// stayed, then the arrivals from row jd + 1 (moving down), each group in its
// source index order.  This is a fixed permutation (the same every run) but
// not the reference's hole-filling order of boundary_periodic.f90:193-226,
// which the FP64 form keeps; cumcnt and np2 mean exactly what they mean
// after sort.f90:36-82.

__global__ void k_fill_int(int *p, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = 0;
}

// One workgroup per destination segment: cumcnt, np2 and, in place of the
// three source counts that land here, their first destination slot.
__global__ __launch_bounds__(256) void k_bin_prefix(int nx, int ny, int cap, int *__restrict__ bcnt,
                                                    int *__restrict__ cumcnt,
                                                    int *__restrict__ np2new,
                                                    int *__restrict__ derr) {
  __shared__ int sh[256];
  const int dseg = blockIdx.x;
  const int s = dseg / ny, jd = dseg % ny;
  const int jlo = jd == 0 ? ny - 1 : jd - 1, jhi = jd == ny - 1 ? 0 : jd + 1;
  int *lo = bcnt + (((size_t)s * ny + jlo) * 3 + 2) * nx;   // row jd-1 moving up
  int *mid = bcnt + (((size_t)s * ny + jd) * 3 + 1) * nx;   // staying
  int *hi = bcnt + (((size_t)s * ny + jhi) * 3 + 0) * nx;   // row jd+1 moving down
  const int per = (nx + 255) / 256;
  const int c0 = min(nx, (int)threadIdx.x * per), c1 = min(nx, c0 + per);
  int tot = 0;
  for (int c = c0; c < c1; ++c) tot += lo[c] + mid[c] + hi[c];
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - tot;
  int *cc = cumcnt + (size_t)dseg * (nx + 1);
  for (int c = c0; c < c1; ++c) {
    const int a = lo[c], b = mid[c], d = hi[c];
    cc[c] = run;
    lo[c] = run;
    mid[c] = run + a;
    hi[c] = run + a + b;
    run += a + b + d;
  }
  if (threadIdx.x == 255) {
    cc[nx] = sh[255];
    np2new[dseg] = sh[255];
    if (sh[255] > cap) atomicOr(derr, DERR_NP_OVER);
  }
}

// Stable scatter, one wavefront per (source segment, range of SUBC old
// cells).  The source is sorted by old cell, and a particle lands at most one
// cell away, so a range's particles can only share destination keys with the
// old cells just outside it.  Those that come earlier in the source order
// (old cells below c0, including the periodic wrap of the last range) are
// counted first; the range then writes its particles in order behind them.
constexpr int SUBC = 32;

// Local key of landing cell k (-1 if outside the range's key set): the
// range's own cells, then the periodic neighbours c0 - 1 and c1.
__device__ inline int local_key(int k, int c0, int c1, int nx, int nk) {
  if (k >= c0 && k < c1) return k - c0 + 1;
  if (k == (c0 == 0 ? nx - 1 : c0 - 1)) return 0;
  if (k == (c1 == nx ? 0 : c1)) return nk - 1;
  return -1;
}

__device__ inline int row_move(float oy) { return oy < 0.0f ? -1 : (oy >= 1.0f ? 1 : 0); }

constexpr int SCW = 4;     // wavefronts (ranges) per workgroup

__global__ __launch_bounds__(SW * SCW) void k_bin_scatter(SP P, int ny, int nsub,
                                                          const int *__restrict__ ccold,
                                                          Rec32 src, const int64_t *__restrict__ ids,
                                                          Rec32 dst, int64_t *__restrict__ idd,
                                                          const int *__restrict__ bcnt,
                                                          int *__restrict__ derr) {
  __shared__ int offs[SCW][3 * (SUBC + 2)];
  __shared__ int preds[SCW][8];
  __shared__ int npreds[SCW];
  const int wv = threadIdx.x / SW;
  const int range = blockIdx.x * SCW + wv;
  if (range >= nsub) return;                       // whole wavefronts only: no barrier below
  int *off = offs[wv], *pred = preds[wv];
  int &npred = npreds[wv];
  const int seg = blockIdx.y;
  const int s = seg / ny, jl = seg % ny;
  const int lane = threadIdx.x % SW;
  const int nx = P.nx;
  const int c0 = range * SUBC, c1 = min(nx, c0 + SUBC);
  const bool whole = c0 == 0 && c1 == nx;           // one range: no neighbour keys
  const int nk = c1 - c0 + 2;
  const int *cc = ccold + (size_t)seg * (nx + 1);
  const size_t base = (size_t)seg * P.cap;
  // first destination slots of the range's keys
  for (int e = lane; e < 3 * nk; e += SW) {
    const int d = e / nk, kl = e % nk;
    int cell = c0 - 1 + kl;
    if (kl == 0) cell = c0 == 0 ? nx - 1 : c0 - 1;
    if (kl == nk - 1) cell = c1 == nx ? 0 : c1;
    off[e] = (whole && (kl == 0 || kl == nk - 1)) ? 0
                                                   : bcnt[((size_t)seg * 3 + d) * nx + cell];
  }
  // old cells before c0 (in the source order) that neighbour one of the
  // range's keys; only the edge keys c0 - 1, c0, c1 - 1, c1 (periodic) have
  // neighbours outside [c0, c1), and these are among c0 - 2, c0 - 1, 0, 1
  if (lane == 0) {
    int np = 0;
    if (!whole) {
      const int edge[4] = {c0 == 0 ? nx - 1 : c0 - 1, c0, c1 - 1, c1 == nx ? 0 : c1};
      for (int q = 0; q < 4; ++q) {
        for (int d = -1; d <= 1; ++d) {
          int o = edge[q] + d;
          o = o < 0 ? o + nx : (o >= nx ? o - nx : o);
          if (o >= c0) continue;
          bool dup = false;
          for (int t = 0; t < np; ++t) dup = dup || pred[t] == o;
          if (!dup && np < 8) pred[np++] = o;
        }
      }
    }
    npred = np;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int q = 0; q < npred; ++q) {
    const int o = pred[q];
    for (int k = cc[o] + lane; k < cc[o + 1]; k += SW) {
      const size_t p = base + k;
      const int kl = local_key(src.ci[p], c0, c1, nx, nk);
      if (kl >= 0) atomicAdd(&off[(row_move(src.oy[p]) + 1) * nk + kl], 1);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int bad = 0;
  const int kb = cc[c0], ke = cc[c1];
  for (int k0 = kb; k0 < ke; k0 += SW * BT) {
    int key[BT], dy[BT];
    Rec32::Val v[BT];
    int64_t vid[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = k0 + b * SW + lane;
      key[b] = -1;
      dy[b] = 0;
      if (k < ke) {
        const size_t p = base + k;
        v[b] = src.load(p);
        vid[b] = ids[p];
        dy[b] = row_move(v[b].oy);
        const int kl = (v[b].ci >= 0 && v[b].ci < nx) ? local_key(v[b].ci, c0, c1, nx, nk) : -1;
        key[b] = kl >= 0 ? (dy[b] + 1) * nk + kl : -1;
        if (key[b] < 0) bad = 1;
      }
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const bool act = key[b] >= 0;
      const WaveMatch m = wave_match(key[b], act);
      if (act) {
        const int slot = off[key[b]] + m.rank;
        const int jd = jl + dy[b] < 0 ? ny - 1 : (jl + dy[b] >= ny ? 0 : jl + dy[b]);
        if (slot < P.cap) {
          Rec32::Val w = v[b];
          w.oy = dy[b] > 0 ? w.oy - 1.0f : (dy[b] < 0 ? wm_below_one(w.oy + 1.0f) : w.oy);
          const size_t d = ((size_t)s * ny + jd) * P.cap + slot;
          dst.store(d, w);
          idd[d] = vid[b];
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (act && m.leader) off[key[b]] += m.count;
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (bad) atomicOr(derr, DERR_RANGE);
}

}  // namespace

int wm_k_bin_clear(wm_ctx *ctx) {
  const size_t n = (size_t)ctx->g.nsp * ctx->g.ny * 3 * ctx->g.nx;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)),
                     dim3(256), 0, ctx->stream, ctx->bcnt, n);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_bin32(wm_ctx *ctx, int src, int dst) {
  SP P;
  P.nx = ctx->g.nx;
  P.nxgs = ctx->g.nxgs;
  P.cap = ctx->cap;
  const int ny = ctx->g.ny;
  const int nseg = ctx->g.nsp * ny;
  // the new np2 and cumcnt go to spare arrays, which then take their place
  // (the scatter reads the old cumcnt)
  hipLaunchKernelGGL(k_bin_prefix, dim3(nseg), dim3(256), 0, ctx->stream, P.nx, ny, P.cap,
                     ctx->bcnt, ctx->cumcnt2, ctx->nup, ctx->derr);
  const int nsub = wm_cdiv(P.nx, SUBC);
  hipLaunchKernelGGL(k_bin_scatter, dim3(wm_cdiv(nsub, SCW), nseg), dim3(SW * SCW), 0, ctx->stream,
                     P, ny, nsub, ctx->cumcnt, ctx->b32[src], ctx->id[ctx->idcur], ctx->b32[dst],
                     ctx->id[1 - ctx->idcur], ctx->bcnt, ctx->derr);
  WM_HIP(hipGetLastError());
  std::swap(ctx->np2, ctx->nup);
  std::swap(ctx->cumcnt, ctx->cumcnt2);
  ctx->idcur = 1 - ctx->idcur;
  return WM_OK;
}

int wm_k_sort(wm_ctx *ctx, int src, int dst) {
  SP P;
  P.nx = ctx->g.nx;
  P.nxgs = ctx->g.nxgs;
  P.cap = ctx->cap;
  const int nseg = ctx->g.nsp * ctx->g.ny;
  const size_t lds = sizeof(int) * (size_t)P.nx;
  if (lds > 65536) {
    wm_set_error(ctx, WM_ERR_ARG, "sort: nx=%d exceeds the 16384 cells one LDS histogram holds", P.nx);
    return WM_ERR_ARG;
  }
  if (ctx->form == WM_FORM_FP32) {
    hipLaunchKernelGGL(k_sort<Rec32>, dim3(nseg), dim3(SW), lds, ctx->stream, P, ctx->np2,
                       ctx->b32[src], ctx->id[ctx->idcur], ctx->b32[dst], ctx->id[1 - ctx->idcur],
                       ctx->cumcnt, ctx->derr);
  } else {
    hipLaunchKernelGGL(k_sort<Rec64>, dim3(nseg), dim3(SW), lds, ctx->stream, P, ctx->np2,
                       ctx->b64(src), ctx->id[ctx->idcur], ctx->b64(dst), ctx->id[1 - ctx->idcur],
                       ctx->cumcnt, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  ctx->idcur = 1 - ctx->idcur;
  return WM_OK;
}
