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

}  // namespace

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
