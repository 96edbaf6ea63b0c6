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
#include "wm_migrate.h"

namespace {

using namespace wm_mig;

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
// FP32 form: cell bins.  Bin (s, jl, i) = seg * nx + i holds up to CB
// particles of cell i of row segment seg = s * ny + jl in slots
// [bin * CB, bin * CB + cnt[bin]).

// Dense rows (any order, np2 per segment) -> bins, stable: within a bin the
// particles keep their row order.  One wavefront per segment, as k_sort.
__global__ __launch_bounds__(SW) void k_dense_to_bins(SP P, int CB, const int *__restrict__ np2,
                                                      Rec32 src, const int64_t *__restrict__ ids,
                                                      Bins32 dst, int64_t *__restrict__ idd,
                                                      int *__restrict__ cnt, int *derr) {
  extern __shared__ int run[];             // particles written per cell so far
  const int seg = blockIdx.x;
  const int lane = threadIdx.x;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  int bad = 0;
  for (int i = lane; i < P.nx; i += SW) run[i] = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += SW) {
    const int k = k0 + lane;
    Rec32::Val v{};
    int key = -1;
    int64_t vid = 0;
    if (k < n) {
      v = src.load(base + k);
      vid = ids[base + k];
      key = v.ci;
      if (key < 0 || key >= P.nx) {
        bad |= DERR_RANGE;
        key = -1;
      }
    }
    const bool act = key >= 0;
    const WaveMatch m = wave_match(key, act);
    if (act) {
      const int slot = run[key] + m.rank;
      if (slot < CB) {
        const size_t d = ((size_t)seg * P.nx + key) * CB + slot;
        dst.q[d] = make_float4(v.ox, v.oy, v.ux, v.uy);
        dst.uz[d] = v.uz;
        idd[d] = vid;
      } else {
        bad |= DERR_NP_OVER;
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (act && m.leader) run[key] += m.count;
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (int i = lane; i < P.nx; i += SW) cnt[(size_t)seg * P.nx + i] = min(run[i], CB);
  if (bad) atomicOr(derr, bad);
}

// Bins -> dense rows sorted by cell, with cumcnt and np2 as sort.f90:64-69
// leaves them.  One workgroup per segment.
__global__ __launch_bounds__(256) void k_bins_to_dense(SP P, int CB, const int *__restrict__ cnt,
                                                       Bins32 src, const int64_t *__restrict__ ids,
                                                       Rec32 dst, int64_t *__restrict__ idd,
                                                       int *__restrict__ cumcnt,
                                                       int *__restrict__ np2, int *derr) {
  extern __shared__ int cc[];              // nx + 1 prefix
  __shared__ int sh[256];
  const int seg = blockIdx.x;
  const int *c = cnt + (size_t)seg * P.nx;
  const int per = (P.nx + 255) / 256;
  const int c0 = min(P.nx, (int)threadIdx.x * per), c1 = min(P.nx, c0 + per);
  int tot = 0;
  for (int i = c0; i < c1; ++i) tot += c[i];
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - tot;
  for (int i = c0; i < c1; ++i) {
    cc[i] = run;
    run += c[i];
  }
  if (threadIdx.x == 255) cc[P.nx] = sh[255];
  __syncthreads();
  const int n = cc[P.nx];
  int *cco = cumcnt + (size_t)seg * (P.nx + 1);
  for (int i = threadIdx.x; i <= P.nx; i += 256) cco[i] = cc[i];
  if (threadIdx.x == 0) {
    np2[seg] = n;
    if (n > P.cap) atomicOr(derr, DERR_NP_OVER);
  }
  const size_t base = (size_t)seg * P.cap;
  for (int k = threadIdx.x; k < min(n, P.cap); k += 256) {
    int lo = 0, hi = P.nx - 1;                     // cell i with cc[i] <= k < cc[i + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (cc[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const size_t f = ((size_t)seg * P.nx + lo) * CB + (k - cc[lo]);
    const size_t d = base + k;
    const float4 v = src.q[f];
    dst.ci[d] = lo;
    dst.ox[d] = v.x;
    dst.oy[d] = v.y;
    dst.ux[d] = v.z;
    dst.uy[d] = v.w;
    dst.uz[d] = src.uz[f];
    idd[d] = ids[f];
  }
}

// The settle kernel: the migration of every bin of rows r0 .. r0 + nr - 1 from
// the last push's summaries and records, when no push follows to do it in
// its prologue (state download, energy sums).
__global__ void k_migrate(Mig G, Movers M, Bins32 A, int64_t *__restrict__ ids, int *derr) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t / DL < G.nbin;
       t += (size_t)gridDim.x * blockDim.x) {
    // cell (i, jl) and species sp of the group's bin (bins fit 32 bits)
    const unsigned du = (unsigned)(t / DL);
    const int i = (int)(du % (unsigned)G.nx);
    const unsigned sj = du / (unsigned)G.nx;
    const int jl = G.r0 + (int)(sj % (unsigned)G.nr), sp = (int)(sj / (unsigned)G.nr);
    const size_t d = ((size_t)sp * G.ny + jl) * G.nx + i;
    (void)migrate_bin(G, M, A, ids, derr, d, i, jl, sp, (int)(t % DL));
  }
}

}  // namespace

int wm_k_dense_to_bins(wm_ctx *ctx) {
  SP P;
  P.nx = ctx->g.nx;
  P.nxgs = ctx->g.nxgs;
  P.cap = ctx->cap;
  const int nseg = ctx->g.nsp * ctx->g.ny;
  hipLaunchKernelGGL(k_dense_to_bins, dim3(nseg), dim3(SW), sizeof(int) * P.nx, ctx->stream, P,
                     ctx->CB, ctx->np2, ctx->b32[1], ctx->id[ctx->idcur], ctx->bins,
                     ctx->id_bins, ctx->bin_cnt, ctx->derr);
  WM_HIP(hipGetLastError());
  ctx->mv_pending = false;                         // the bins are the whole state now
  return WM_OK;
}

int wm_k_bins_to_dense(wm_ctx *ctx) {
  WM_TRY(wm_settle_movers(ctx));
  SP P;
  P.nx = ctx->g.nx;
  P.nxgs = ctx->g.nxgs;
  P.cap = ctx->cap;
  const int nseg = ctx->g.nsp * ctx->g.ny;
  hipLaunchKernelGGL(k_bins_to_dense, dim3(nseg), dim3(256), sizeof(int) * (P.nx + 1),
                     ctx->stream, P, ctx->CB, ctx->bin_cnt, ctx->bins, ctx->id_bins,
                     ctx->b32[1], ctx->id[ctx->idcur], ctx->cumcnt, ctx->np2, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_settle_movers(wm_ctx *ctx) {
  if (ctx->form != WM_FORM_FP32 || !ctx->mv_pending) return WM_OK;
  Mig G;
  G.nx = ctx->g.nx;
  G.ny = ctx->g.ny;
  G.nxr = (int)((size_t)ctx->nregion / (size_t)ctx->g.ny);
  G.CB = ctx->CB;
  G.r0 = 0;
  G.nr = ctx->g.ny;
  G.nbin = (size_t)ctx->g.nsp * G.ny * G.nx;
  G.sum = ctx->bin_sum[ctx->mv_last];
  G.cnt = ctx->bin_cnt;
  hipLaunchKernelGGL(k_migrate, dim3((unsigned)std::min<size_t>((G.nbin * DL + 255) / 256, 65536)),
                     dim3(256), 0, ctx->stream, G, ctx->mv[ctx->mv_last], ctx->bins, ctx->id_bins,
                     ctx->derr);
  WM_HIP(hipGetLastError());
  ctx->mv_pending = false;
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
    // the FP32 form keeps its particles in cell bins (wm_k_dense_to_bins)
    wm_set_error(ctx, WM_ERR_ARG, "sort__bucket works on the FP64 particle form");
    return WM_ERR_ARG;
  }
  {
    hipLaunchKernelGGL(k_sort<Rec64>, dim3(nseg), dim3(SW), lds, ctx->stream, P, ctx->np2,
                       ctx->b64(src), ctx->id[ctx->idcur], ctx->b64(dst), ctx->id[1 - ctx->idcur],
                       ctx->cumcnt, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  ctx->idcur = 1 - ctx->idcur;
  return WM_OK;
}
