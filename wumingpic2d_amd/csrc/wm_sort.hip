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

struct Mig {                    // bins and mover regions of one migration
  int nx, ny, nxr, CB;          // nxr: regions per row (cells rounded up to the push chunk)
  int r0, nr;                   // the migrated rows (every species)
  size_t nbin;                  // bins of those rows
  const BinSum *sum;            // per bin: the push's summary of the movers that left it
  int *cnt;
};

// First record of the movers of the bin in cell (i, jl) in the cell's region
// (out: the bin summary's (offset << 8) | count).
__device__ inline size_t bin_records(const Mig &G, int i, int jl, uint32_t out) {
  return ((size_t)jl * G.nxr + i) * WM_MG + (out >> 8);
}

// Byte k of a summary's direction counts (words lo, hi), and entry j of its
// mover list (words w0..w3).  The words are values, not a dynamically indexed
// array, which the compiler would place in memory.
__device__ inline int sum_cnt(uint32_t lo, uint32_t hi, int k) {
  const uint32_t v = k < 4 ? lo : hi;
  return (int)((v >> (8 * (k & 3))) & 0xFFu);
}
__device__ inline int sum_mv(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int j) {
  const uint32_t v = j < 4 ? (j < 2 ? w0 : w1) : (j < 6 ? w2 : w3);
  return (int)((v >> (16 * (j & 1))) & 0xFFFFu);
}
// Slot of bin d's h-th hole: from its summary for the first WM_SUM_MV, else
// from the record.
__device__ inline int hole_slot(const Movers &M, size_t r0, uint4 l, int h) {
  return h < WM_SUM_MV ? sum_mv(l.x, l.y, l.z, l.w, h) & ((1 << WM_SLOT_BITS) - 1)
                       : M.rec[r0 + h].sd & 0xFFFFFF;
}

// The migration of bin d, by a group of 8 lanes that alone writes d (no
// lane reads what another group writes).  Lane k takes the arrivals from the
// source bin in direction index k (directions 0, 1, 2, 3, 5, 6, 7, 8), in the
// order the source listed them; directions follow one another, so arrival
// index = the earlier directions' count (a prefix over the group's lanes) +
// the rank within direction k.  d held n particles and lost m (its movers'
// slots, ascending, are its holes): the a-th arrival takes the a-th hole
// while there are holes, then slot n + (a - m).  With fewer arrivals than
// holes, lane 0 fills the holes left below the new count, in order, with the
// particles above it.  The summaries of d and of the source (two 32-byte
// loads, issued together) give the counts and the first WM_SUM_MV movers'
// slots and directions; only a bin with more movers reads the rest from its
// records.
constexpr int DL = 8;
__global__ void k_migrate(Mig G, Movers M, Bins32 A, int64_t *__restrict__ ids, int *derr) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t / DL < G.nbin;
       t += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % DL);
    const int dir = k < 4 ? k : k + 1;
    // cell (i, jl) and species sp of the group's bin d (bins fit 32 bits),
    // and its source
    const unsigned du = (unsigned)(t / DL);
    const int i = (int)(du % (unsigned)G.nx);
    const unsigned sj = du / (unsigned)G.nx;
    const int jl = G.r0 + (int)(sj % (unsigned)G.nr), sp = (int)(sj / (unsigned)G.nr);
    const size_t d = ((size_t)sp * G.ny + jl) * G.nx + i;
    const int si = wm_wrap(i - (dir % 3 - 1), G.nx), sjl = wm_wrap(jl - (dir / 3 - 1), G.ny);
    const size_t src = ((size_t)sp * G.ny + sjl) * G.nx + si;
    const uint4 *own = (const uint4 *)(G.sum + d), *from = (const uint4 *)(G.sum + src);
    uint4 oh = own[0], ol = own[1], fh = from[0], fl = from[1];
    int nd = G.cnt[d];
    // all loads in flight together (the compiler would otherwise issue some
    // only where they are first used, a round trip later)
    __asm__ volatile("" : "+v"(ol.x), "+v"(ol.y), "+v"(ol.z), "+v"(ol.w), "+v"(fh.x), "+v"(fh.y),
                     "+v"(nd));
    const int n = min(nd, G.CB);
    const int c = sum_cnt(fh.x, fh.y, k);
    const int m = (int)(oh.z & 0xFFu);
    const size_t r0 = bin_records(G, i, jl, oh.z);
    int inc = c;                                   // inclusive prefix over the group
#pragma unroll
    for (int o = 1; o < DL; o <<= 1) {
      const int v = __shfl_up(inc, o, DL);
      if (k >= o) inc += v;
    }
    const int total = __shfl(inc, DL - 1, DL);
    if (c > 0) {
      const int ms = (int)(fh.z & 0xFFu);
      const size_t rs = bin_records(G, si, sjl, fh.z);
      int a = inc - c, left = c;
      for (int j = 0; j < ms && left > 0; ++j) {
        const int kj = j < WM_SUM_MV ? sum_mv(fl.x, fl.y, fl.z, fl.w, j) >> WM_SLOT_BITS
                                     : M.rec[rs + j].sd >> 24;
        if (kj != k) continue;
        const MoverRec mr = M.rec[rs + j];
        --left;
        const int slot = a < m ? hole_slot(M, r0, ol, a) : n + (a - m);
        ++a;
        if (slot >= G.CB) {
          atomicOr(derr, DERR_NP_OVER);
          continue;
        }
        const size_t p = d * G.CB + slot;
        A.q[p] = make_float4(mr.ox1, mr.oy1, mr.ux, mr.uy);
        A.uz[p] = mr.uz;
        ids[p] = mr.id;
      }
    }
    if (k != 0) continue;
    const int n1 = n - m;
    if (total < m) {
      const size_t base = d * G.CB;
      const int nnew = n1 + total;
      int hb = total;                              // holes below the new count
      while (hb < m && hole_slot(M, r0, ol, hb) < nnew) ++hb;
      int q = total, h = hb;
      for (int from_slot = nnew; from_slot < n && q < hb; ++from_slot) {
        if (h < m && hole_slot(M, r0, ol, h) == from_slot) {       // a hole above the new count
          ++h;
          continue;
        }
        const size_t dst = base + hole_slot(M, r0, ol, q), sf = base + from_slot;
        A.q[dst] = A.q[sf];
        A.uz[dst] = A.uz[sf];
        ids[dst] = ids[sf];
        ++q;
      }
    }
    int cn = n1 + total;
    if (cn > G.CB) {
      atomicOr(derr, DERR_NP_OVER);
      cn = G.CB;
    }
    G.cnt[d] = cn;
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
  return WM_OK;
}

int wm_k_bins_to_dense(wm_ctx *ctx) {
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

int wm_k_insert_movers(wm_ctx *ctx, int r0, int nrows) {
  if (nrows < 0) nrows = ctx->g.ny - r0;
  if (r0 < 0 || nrows < 1 || r0 + nrows > ctx->g.ny) {
    wm_set_error(ctx, WM_ERR_ARG, "migration: rows %d..%d outside 0..%d", r0, r0 + nrows,
                 ctx->g.ny);
    return WM_ERR_ARG;
  }
  Mig G;
  G.nx = ctx->g.nx;
  G.ny = ctx->g.ny;
  G.nxr = (int)((size_t)ctx->nregion / (size_t)ctx->g.ny);
  G.CB = ctx->CB;
  G.r0 = r0;
  G.nr = nrows;
  G.nbin = (size_t)ctx->g.nsp * nrows * G.nx;
  G.sum = ctx->bin_sum;
  G.cnt = ctx->bin_cnt;
  // a group of DL lanes never straddles a wave (the stride is a whole number
  // of waves), and the loop bound is per group, so the shuffles see all lanes
  hipLaunchKernelGGL(k_migrate,
                     dim3((unsigned)std::min<size_t>((G.nbin * DL + 255) / 256, 65536)), dim3(256),
                     0, ctx->stream, G, ctx->mv, ctx->bins, ctx->id_bins, ctx->derr);
  WM_HIP(hipGetLastError());
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
