// wm_sort.hip -- stable per-row bucket sort by x-cell (sort.f90:36-82).
//
// One workgroup per (species, row) segment.  The cell histogram lives in LDS
// (nx counters); counting and scattering walk the segment in index order in
// tiles of 256.  Inside a wavefront, lanes with equal keys are grouped by a
// ballot loop over the distinct keys (sorted input has one to three per
// wavefront), which gives each lane its rank among equal keys below it; the
// four wavefronts of a tile exchange their (key, count) lists through LDS so
// that the destination of every particle is
//   start[key] + (equal keys in earlier tiles and wavefronts) + rank,
// i.e. exactly the reference's stable scatter.  Keys are int(x) as in
// sort.f90:60 (the reference assumes delx = 1).
#include "wm_internal.h"

namespace {

constexpr int SB = 256;
constexpr int NW = SB / 64;

struct SP {
  int nx, nxgs, cap;
};

struct WaveMatch {
  int rank;     // equal keys in lower lanes
  int count;    // equal keys in the wavefront
  int slot;     // index of this key in the wavefront's distinct-key list
  bool leader;  // lowest lane holding the key
  int ndist;    // distinct keys in the wavefront
};

__device__ inline WaveMatch wave_match(int key, bool active) {
  const int lane = threadIdx.x & 63;
  WaveMatch w{0, 0, 0, false, 0};
  unsigned long long rem = __ballot(active);
  int it = 0;
  while (rem) {
    const int l = __ffsll((long long)rem) - 1;
    const int k = __shfl(key, l);
    const unsigned long long m = __ballot(active && key == k);
    if (active && key == k) {
      w.rank = __popcll(m & ((1ULL << lane) - 1ULL));
      w.count = __popcll(m);
      w.slot = it;
      w.leader = (lane == l);
    }
    rem &= ~m;
    ++it;
  }
  w.ndist = it;
  return w;
}

__global__ __launch_bounds__(SB) void k_sort(SP P, const int *__restrict__ np2,
                                             const double *__restrict__ xs,
                                             const double *__restrict__ ys,
                                             const double *__restrict__ uxs,
                                             const double *__restrict__ uys,
                                             const double *__restrict__ uzs,
                                             const int64_t *__restrict__ ids,
                                             double *__restrict__ xd, double *__restrict__ yd,
                                             double *__restrict__ uxd, double *__restrict__ uyd,
                                             double *__restrict__ uzd, int64_t *__restrict__ idd,
                                             int *__restrict__ cumcnt, int *derr) {
  extern __shared__ int cnt[];             // nx counters, then offsets
  __shared__ int wl_key[NW][64], wl_cnt[NW][64], wl_n[NW];
  __shared__ int scan_sh[SB];
  const int seg = blockIdx.x;
  const int n = np2[seg];
  const size_t base = (size_t)seg * P.cap;
  const int w = threadIdx.x >> 6;
  int bad = 0;
  for (int i = threadIdx.x; i < P.nx; i += SB) cnt[i] = 0;
  __syncthreads();
  // count (sort.f90:59-62)
  for (int k0 = 0; k0 < n; k0 += SB) {
    const int k = k0 + threadIdx.x;
    int key = 0;
    bool act = k < n;
    if (act) {
      key = (int)xs[base + k] - P.nxgs;
      if (key < 0 || key >= P.nx) {
        bad = 1;
        act = false;
      }
    }
    const WaveMatch m = wave_match(key, act);
    if (act && m.leader) atomicAdd(&cnt[key], m.count);
  }
  __syncthreads();
  // exclusive prefix -> cumcnt(nxs:nxe+1) (sort.f90:64-69); cnt becomes start
  {
    const int per = (P.nx + SB - 1) / SB;
    const int lo = min(P.nx, threadIdx.x * per), hi = min(P.nx, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    scan_sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SB; off <<= 1) {
      const int v = (int)threadIdx.x >= off ? scan_sh[threadIdx.x - off] : 0;
      __syncthreads();
      scan_sh[threadIdx.x] += v;
      __syncthreads();
    }
    int run = scan_sh[threadIdx.x] - s;     // exclusive
    int *cc = cumcnt + (size_t)seg * (P.nx + 1);
    for (int i = lo; i < hi; ++i) {
      const int c = cnt[i];
      cc[i] = run;
      cnt[i] = run;
      run += c;
    }
    if (threadIdx.x == SB - 1) cc[P.nx] = scan_sh[SB - 1];
  }
  __syncthreads();
  // stable scatter (sort.f90:71-75)
  for (int k0 = 0; k0 < n; k0 += SB) {
    const int k = k0 + threadIdx.x;
    int key = 0;
    bool act = k < n;
    if (act) {
      key = (int)xs[base + k] - P.nxgs;
      act = key >= 0 && key < P.nx;
    }
    const WaveMatch m = wave_match(key, act);
    if (act && m.leader) {
      wl_key[w][m.slot] = key;
      wl_cnt[w][m.slot] = m.count;
    }
    if ((threadIdx.x & 63) == 0) wl_n[w] = m.ndist;
    __syncthreads();
    if (act) {
      int off = cnt[key] + m.rank;
      for (int v = 0; v < w; ++v)
        for (int e = 0; e < wl_n[v]; ++e) off += (wl_key[v][e] == key) ? wl_cnt[v][e] : 0;
      const size_t s = base + k, d = base + off;
      xd[d] = xs[s];
      yd[d] = ys[s];
      uxd[d] = uxs[s];
      uyd[d] = uys[s];
      uzd[d] = uzs[s];
      idd[d] = ids[s];
    }
    __syncthreads();
    if (act && m.leader) atomicAdd(&cnt[key], m.count);
    __syncthreads();
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
  const PBank &S = ctx->bank[src];
  const PBank &D = ctx->bank[dst];
  hipLaunchKernelGGL(k_sort, dim3(nseg), dim3(SB), lds, ctx->stream, P, ctx->np2, S.x, S.y, S.ux,
                     S.uy, S.uz, ctx->id[ctx->idcur], D.x, D.y, D.ux, D.uy, D.uz,
                     ctx->id[1 - ctx->idcur], ctx->cumcnt, ctx->derr);
  WM_HIP(hipGetLastError());
  ctx->idcur = 1 - ctx->idcur;
  return WM_OK;
}
