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
#include "wm_bins.h"
#include "wm_migrate.h"

namespace {

using namespace wm_mig;
using wm_bins::WaveMatch;
using wm_bins::wave_match;

constexpr int SW = 64;     // one wavefront per segment
constexpr int BT = 4;      // tiles in flight

struct SP {
  int nx, nxgs, cap;
};

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
// FP32 form: cell bins inside the row segments.  Bin (seg, i) occupies slots
// [bstart[seg][i], bstart[seg][i + 1]) of segment seg (from seg * cap) and
// holds cnt[seg * nx + i] particles at its front.

struct SB {
  int nx, ny, nxgs, nys, cap, nsp;
};

__device__ inline size_t seg_at(const SB &P, int seg) { return (size_t)seg * P.cap; }

// Staged reference-layout rows -> bins, stable within a bin (row order).
// Segment seg0 + b holds cnt[b] records of ndim doubles at st + (b * w + k) *
// ndim.  One wavefront per segment: the cell histogram, the segment's layout
// (wm_bins::wave_layout), then the scatter in row order.  ci = int(x) - nxgs,
// ox = x - int(x) rounded to FP32 and kept below 1, oy likewise against the
// row j.
__global__ __launch_bounds__(SW) void k_stage_to_bins(SB P, int seg0, int w, int ndim,
                                                      const int *__restrict__ cnt,
                                                      const double *__restrict__ st, Bins32 A,
                                                      int64_t *__restrict__ ids,
                                                      int *__restrict__ bstart,
                                                      int *__restrict__ bin_cnt, int *derr) {
  extern __shared__ int run[];             // particles per cell, then written per cell
  const int b = blockIdx.x, seg = seg0 + b, lane = threadIdx.x;
  const int n = cnt[b];
  const int j = P.nys + seg % P.ny;
  const double *rows = st + (size_t)b * w * ndim;
  int *bs = bstart + (size_t)seg * (P.nx + 1);
  int bad = 0;
  for (int i = lane; i < P.nx; i += SW) run[i] = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += SW) {
    const int k = k0 + lane;
    int key = k < n ? (int)rows[(size_t)k * ndim] - P.nxgs : -1;
    if (k < n && (key < 0 || key >= P.nx)) {
      bad |= DERR_RANGE;
      key = -1;
    }
    const WaveMatch m = wave_match(key, key >= 0);
    if (key >= 0 && m.leader) run[key] += m.count;
  }
  __syncthreads();
  if (!wm_bins::wave_layout([&](int i) { return run[i]; }, P.nx, P.cap, bs)) {
    for (int i = lane; i <= P.nx; i += SW) bs[i] = 0;    // an empty, consistent layout
    for (int i = lane; i < P.nx; i += SW) bin_cnt[(size_t)seg * P.nx + i] = 0;
    if (lane == 0) atomicOr(derr, DERR_NP_OVER | bad);
    return;
  }
  for (int i = lane; i < P.nx; i += SW) {
    bin_cnt[(size_t)seg * P.nx + i] = run[i];
    run[i] = 0;
  }
  __syncthreads();                         // the layout's stores before its loads below
  for (int k0 = 0; k0 < n; k0 += SW) {
    const int k = k0 + lane;
    const double *r = rows + (size_t)min(k, n - 1) * ndim;
    const double x = r[0], y = r[1];
    const int ix = (int)x;
    const bool act = k < n && ix - P.nxgs >= 0 && ix - P.nxgs < P.nx;
    const int key = act ? ix - P.nxgs : -1;
    const WaveMatch m = wave_match(key, act);
    if (act) {
      const size_t p = seg_at(P, seg) + bs[key] + run[key] + m.rank;
      A.q[p] = make_float4(wm_below_one((float)(x - ix)), wm_below_one((float)(y - j)), (float)r[2],
                           (float)r[3]);
      A.uz[p] = (float)r[4];
      if (ids) ids[p] = ndim >= 6 ? __double_as_longlong(r[5]) : 0;
    }
    __builtin_amdgcn_wave_barrier();       // every lane has read its offset before the
    if (act && m.leader) run[key] += m.count;   // leaders advance them
    __builtin_amdgcn_wave_barrier();
  }
  if (bad) atomicOr(derr, bad);
}

// Per segment: cumcnt (exclusive prefix of the bin counts, as sort.f90:64-69
// leaves it) and np2.  One workgroup per segment.
__global__ __launch_bounds__(256) void k_bin_rows(SB P, const int *__restrict__ cnt,
                                                  int *__restrict__ cumcnt, int *__restrict__ np2) {
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
  int *cc = cumcnt + (size_t)seg * (P.nx + 1);
  for (int i = c0; i < c1; ++i) {
    cc[i] = run;
    run += c[i];
  }
  if (threadIdx.x == 255) {
    cc[P.nx] = sh[255];
    np2[seg] = sh[255];
  }
}

// Bins -> staged reference-layout rows sorted by cell (segment seg0 + b at
// st + b * w * ndim), with cumcnt from k_bin_rows: x = nxgs + i + ox and
// y = j + oy in FP64.  One workgroup per segment.
__global__ __launch_bounds__(256) void k_bins_to_stage(SB P, int seg0, int w, int ndim,
                                                       const int *__restrict__ cumcnt,
                                                       const int *__restrict__ bstart, Bins32 A,
                                                       const int64_t *__restrict__ ids,
                                                       double *__restrict__ st, int *derr) {
  extern __shared__ int cc[];              // nx + 1 prefix
  const int b = blockIdx.x, seg = seg0 + b;
  const int j = P.nys + seg % P.ny;
  const int *cg = cumcnt + (size_t)seg * (P.nx + 1);
  for (int i = threadIdx.x; i <= P.nx; i += 256) cc[i] = cg[i];
  __syncthreads();
  const int n = cc[P.nx];
  if (n > w) {
    if (threadIdx.x == 0) atomicOr(derr, DERR_NP_OVER);
    return;
  }
  const int *bs = bstart + (size_t)seg * (P.nx + 1);
  double *rows = st + (size_t)b * w * ndim;
  for (int k = threadIdx.x; k < n; k += 256) {
    int lo = 0, hi = P.nx - 1;                     // cell i with cc[i] <= k < cc[i + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (cc[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const size_t f = seg_at(P, seg) + bs[lo] + (k - cc[lo]);
    const float4 v = A.q[f];
    double *r = rows + (size_t)k * ndim;
    r[0] = (double)(lo + P.nxgs) + (double)v.x;
    r[1] = (double)j + (double)v.y;
    r[2] = v.z;
    r[3] = v.w;
    r[4] = A.uz[f];
    if (ndim >= 6) r[5] = __longlong_as_double(ids ? ids[f] : 0);
  }
}

// After a push: every bin's count once the push's movers are placed
// (count - movers out + arrivals from the 8 neighbours' summaries), into pc.
// A bin whose count would exceed its room flags its row for a re-layout.
// Also keeps the capacity statistics.
__global__ __launch_bounds__(256) void k_fill_check(Mig G, int *__restrict__ pc,
                                                    int *__restrict__ rowflag,
                                                    const int *__restrict__ mfill,
                                                    const int *__restrict__ sfill,
                                                    int *__restrict__ stats, int *derr) {
  const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int cnew = 0, mout = 0;
  bool over = false;
  if (d < G.nbin) {
    const int i = (int)(d % (unsigned)G.nx);
    const unsigned sj = (unsigned)(d / (unsigned)G.nx);
    const int jl = (int)(sj % (unsigned)G.ny), sp = (int)(sj / (unsigned)G.ny);
    mout = (int)(((const uint4 *)(G.sum + d))[0].w);
    int arr = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      int si, sjl;
      const size_t src = source_bin(G, i, jl, sp, k, si, sjl);
      const uint2 c = ((const uint2 *)(G.sum + src))[0];
      arr += sum_cnt(c.x, c.y, k);
    }
    cnew = G.cnt[d] - mout + arr;
    pc[d] = cnew;
    size_t base;
    int capd;
    bin_span(G, sj, i, base, capd);
    over = cnew > capd;
    if (over && !rowflag[jl]) atomicOr(rowflag + jl, 1);
  }
  // maxima: one lane per wave, and an atomic only where it raises the value
  // (atomics on one address from every wave would serialise)
  int a = cnew, m = mout;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a = max(a, __shfl_xor(a, o, 64));
    m = max(m, __shfl_xor(m, o, 64));
  }
  const bool any = __ballot(over) != 0ULL;
  if ((threadIdx.x & 63) == 0) {
    if (a > stats[0]) atomicMax(stats + 0, a);
    if (m > stats[1]) atomicMax(stats + 1, m);
    if (any) atomicOr(derr, DERR_RELAYOUT);
    if (any) atomicAdd(stats + 5, __popcll(__ballot(over)));
  }
  if (d < (size_t)G.ny) {
    if (mfill[d] > stats[3]) atomicMax(stats + 3, mfill[d]);
    if (16 * sfill[d] > stats[4]) atomicMax(stats + 4, 16 * sfill[d]);   // blocks of 16
  }
}

// Re-layout.  A row's bins are laid out anew block by block: a block of RB
// cells with a bin that would overflow gets a new layout inside the slots
// its bins span now, each bin its need plus a share of the block's free
// slots in proportion to its need; when the block's span cannot hold its
// bins' needs the whole segment is laid out anew (wm_bins::wave_layout).
// Every other bin keeps its place and takes its pending movers in place, as
// the next push's prologue would, so the whole row holds its movers after.
constexpr int RB = 64;

// Step 1, one wavefront per segment of the listed rows (batch row b,
// species s: block b * nsp + s): the new first slots into tbs and, per
// block of RB cells, whether its bins move (rblk).
__global__ __launch_bounds__(SW) void k_relayout_plan(SB P, const int *__restrict__ list,
                                                      const int *__restrict__ pc,
                                                      const int *__restrict__ bstart,
                                                      int *__restrict__ tbs, int *__restrict__ rblk,
                                                      int *derr) {
  const int bsg = blockIdx.x, lane = threadIdx.x;
  const int seg = (bsg % P.nsp) * P.ny + list[bsg / P.nsp];
  const int *c = pc + (size_t)seg * P.nx;
  const int *bs = bstart + (size_t)seg * (P.nx + 1);
  const int nblk = (P.nx + RB - 1) / RB;
  int *t = tbs + (size_t)bsg * (P.nx + 1);
  int *rb = rblk + (size_t)bsg * nblk;
  bool full = false;
  for (int blk = 0; blk < nblk && !full; ++blk) {
    const int i0 = blk * RB, i1 = min(P.nx, i0 + RB), i = i0 + lane;
    const bool in = i < i1;
    const int cnt = in ? c[i] : 0, b0 = in ? bs[i] : 0, b1 = in ? bs[i + 1] : 0;
    if (__ballot(in && cnt > b1 - b0) == 0ULL) {             // no bin of the block is full
      if (in) t[i] = b0;
      if (lane == 0) rb[blk] = 0;
      continue;
    }
    const int lo = bs[i0], hi = bs[i1];
    const int need = in ? wm_bins::bin_need(cnt) : 0;
    int S = need;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) S += __shfl_xor(S, o, 64);
    if (S > hi - lo || __ballot(cnt > WM_BIN_MAX) != 0ULL) {
      full = true;                                         // wave-uniform
      break;
    }
    const long long units = (hi - lo - S) / WM_BIN_ALIGN;
    constexpr int top = WM_BIN_MAX / WM_BIN_ALIGN * WM_BIN_ALIGN;
    const int room = in ? (int)min((long long)top, need + units * need / S * WM_BIN_ALIGN) : 0;
    int incl = room;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (in) t[i] = lo + incl - room;
    if (lane == 0) rb[blk] = 1;
  }
  if (!full) {
    if (lane == 0) t[P.nx] = bs[P.nx];
    return;
  }
  for (int blk = lane; blk < nblk; blk += SW) rb[blk] = 1;
  if (!wm_bins::wave_layout([&](int i) { return c[i]; }, P.nx, P.cap, t)) {
    for (int i = lane; i <= P.nx; i += SW) t[i] = 0;       // nothing is placed
    if (lane == 0) atomicOr(derr, DERR_NP_OVER);
  }
}

// Step 2: every bin of the listed rows takes its pending movers, 8 lanes per
// bin: a bin of a re-laid block is placed anew (its particles that stay, then
// its arrivals) into the workspace D at its new place; any other bin migrates
// in place.  The bin's count becomes its new count.
__global__ void k_relayout_move(Mig G, SB P, const int *__restrict__ list,
                                const int *__restrict__ tbs, const int *__restrict__ rblk,
                                Movers M, Bins32 A, int64_t *__restrict__ ids, Bins32 D,
                                int64_t *__restrict__ dids, const int *__restrict__ rowpend,
                                int nb, int *derr) {
  const size_t nbin = (size_t)nb * P.nsp * P.nx;
  const int nblk = (P.nx + RB - 1) / RB;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t / DL < nbin;
       t += (size_t)gridDim.x * blockDim.x) {
    const unsigned bl = (unsigned)(t / DL);
    const int k = (int)(t % DL);
    const int i = (int)(bl % (unsigned)P.nx);
    const int bsg = (int)(bl / (unsigned)P.nx);         // batch segment b * nsp + s
    const int s = bsg % P.nsp, jl = list[bsg / P.nsp];
    const size_t seg = (size_t)s * P.ny + jl;
    const size_t d = seg * P.nx + i;
    const bool pend = rowpend[jl] != 0;
    size_t base;
    int capd;
    bin_span(G, seg, i, base, capd);
    if (rblk[(size_t)bsg * nblk + i / RB]) {
      const int *tb = tbs + (size_t)bsg * (P.nx + 1);
      const size_t dbase = (size_t)bsg * P.cap + tb[i];
      const int cn = place_bin(G, M, A, ids, D, dids, d, i, jl, s, k, base, capd, dbase,
                               tb[i + 1] - tb[i], pend, derr);
      if (k == 0) G.cnt[d] = cn;
    } else if (pend) {
      (void)migrate_bin(G, M, A, ids, derr, d, i, jl, s, k, base, capd);
    }
  }
}

// Step 3: the re-laid bins from the workspace into their new places (a
// wavefront per bin), the segments' new first slots; the rows hold their
// movers now.
__global__ __launch_bounds__(256) void k_relayout_copy(SB P, const int *__restrict__ list,
                                                       const int *__restrict__ tbs,
                                                       const int *__restrict__ rblk, Bins32 D,
                                                       const int64_t *__restrict__ dids, Bins32 A,
                                                       int64_t *__restrict__ ids,
                                                       const int *__restrict__ cnt,
                                                       int *__restrict__ bstart,
                                                       int *__restrict__ rowpend) {
  const int bsg = blockIdx.y;
  const int s = bsg % P.nsp, jl = list[bsg / P.nsp];
  const int seg = s * P.ny + jl;
  const int nblk = (P.nx + RB - 1) / RB;
  const int *tb = tbs + (size_t)bsg * (P.nx + 1);
  const int lane = threadIdx.x & 63;
  for (int i = blockIdx.x * 4 + threadIdx.x / 64; i < P.nx; i += gridDim.x * 4) {
    if (!rblk[(size_t)bsg * nblk + i / RB]) continue;
    const int n = cnt[(size_t)seg * P.nx + i];
    const size_t src = (size_t)bsg * P.cap + tb[i], dst = seg_at(P, seg) + tb[i];
    for (int k = lane; k < n; k += 64) {
      A.q[dst + k] = D.q[src + k];
      A.uz[dst + k] = D.uz[src + k];
      if (ids) ids[dst + k] = dids[src + k];
    }
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i <= P.nx; i += blockDim.x) bstart[(size_t)seg * (P.nx + 1) + i] = tb[i];
    if (threadIdx.x == 0 && s == 0) rowpend[jl] = 0;
  }
}

// The settle kernel: the migration of every bin of rows r0 .. r0 + nr - 1 from
// the last push's summaries and records, when no push follows to do it in
// its prologue (state download, energy sums); rows laid out anew since that
// push hold their movers already.
__global__ void k_migrate(Mig G, Movers M, Bins32 A, int64_t *__restrict__ ids,
                          const int *__restrict__ rowpend, int *derr) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t / DL < G.nbin;
       t += (size_t)gridDim.x * blockDim.x) {
    // cell (i, jl) and species sp of the group's bin (bins fit 32 bits)
    const unsigned du = (unsigned)(t / DL);
    const int i = (int)(du % (unsigned)G.nx);
    const unsigned sj = du / (unsigned)G.nx;
    const int jl = G.r0 + (int)(sj % (unsigned)G.nr), sp = (int)(sj / (unsigned)G.nr);
    if (!rowpend[jl]) continue;
    const size_t seg = (size_t)sp * G.ny + jl;
    const size_t d = seg * G.nx + i;
    size_t base;
    int capd;
    bin_span(G, seg, i, base, capd);
    (void)migrate_bin(G, M, A, ids, derr, d, i, jl, sp, (int)(t % DL), base, capd);
  }
}

__global__ void k_fill_i32(int *a, size_t n, int v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}
}  // namespace

namespace {
SB make_sb(const wm_ctx *ctx) {
  SB P;
  P.nx = ctx->g.nx;
  P.ny = ctx->g.ny;
  P.nxgs = ctx->g.nxgs;
  P.nys = ctx->g.nys;
  P.cap = ctx->cap;
  P.nsp = ctx->g.nsp;
  return P;
}
}  // namespace

wm_mig::Mig wm_make_mig(const wm_ctx *ctx, int buf) {
  Mig G;
  G.nx = ctx->g.nx;
  G.ny = ctx->g.ny;
  G.cap = ctx->cap;
  G.mrow = ctx->mrow;
  G.r0 = 0;
  G.nr = ctx->g.ny;
  G.nbin = (size_t)ctx->g.nsp * G.ny * G.nx;
  G.sum = ctx->bin_sum[buf];
  G.bstart = ctx->bstart;
  G.cnt = ctx->bin_cnt;
  return G;
}

int wm_k_stage_to_bins(wm_ctx *ctx, int seg0, int nseg, int w) {
  const SB P = make_sb(ctx);
  const size_t lds = sizeof(int) * (size_t)P.nx;
  if (lds > 65536) {
    wm_set_error(ctx, WM_ERR_ARG, "FP32 form: nx=%d exceeds the 16384 cells one LDS histogram holds",
                 P.nx);
    return WM_ERR_ARG;
  }
  hipLaunchKernelGGL(k_stage_to_bins, dim3(nseg), dim3(SW), lds, ctx->stream, P, seg0, w,
                     ctx->g.ndim, ctx->cnt_scratch, ctx->stage, ctx->bins, ctx->id_bins,
                     ctx->bstart, ctx->bin_cnt, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_bin_rows(wm_ctx *ctx) {
  const SB P = make_sb(ctx);
  hipLaunchKernelGGL(k_bin_rows, dim3(P.nsp * P.ny), dim3(256), 0, ctx->stream, P, ctx->bin_cnt,
                     ctx->cumcnt, ctx->np2);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_bins_to_stage(wm_ctx *ctx, int seg0, int nseg, int w) {
  const SB P = make_sb(ctx);
  hipLaunchKernelGGL(k_bins_to_stage, dim3(nseg), dim3(256), sizeof(int) * (size_t)(P.nx + 1),
                     ctx->stream, P, seg0, w, ctx->g.ndim, ctx->cumcnt, ctx->bstart, ctx->bins,
                     ctx->id_bins, ctx->stage, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_settle_movers(wm_ctx *ctx) {
  if (ctx->form != WM_FORM_FP32 || !ctx->mv_pending) return WM_OK;
  const Mig G = wm_make_mig(ctx, ctx->mv_last);
  hipLaunchKernelGGL(k_migrate, dim3((unsigned)std::min<size_t>((G.nbin * DL + 255) / 256, 65536)),
                     dim3(256), 0, ctx->stream, G, ctx->mv[ctx->mv_last], ctx->bins, ctx->id_bins,
                     ctx->rowpend, ctx->derr);
  hipLaunchKernelGGL(k_fill_i32, dim3(wm_cdiv(ctx->g.ny, 256)), dim3(256), 0, ctx->stream,
                     ctx->rowpend, (size_t)ctx->g.ny, 0);
  WM_HIP(hipGetLastError());
  ctx->mv_pending = false;
  return WM_OK;
}

int wm_k_fill_check(wm_ctx *ctx) {
  const Mig G = wm_make_mig(ctx, ctx->mv_last);
  const size_t nthr = std::max(G.nbin, (size_t)ctx->g.ny);
  hipLaunchKernelGGL(k_fill_check, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, ctx->stream,
                     G, ctx->pcnt, ctx->rowflag, ctx->mfill[ctx->mv_last], ctx->sfill, ctx->fstat,
                     ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// The rows the fill check flagged, laid out anew in batches of relay_rows
// rows through the workspace; their pending movers are placed on the way.
int wm_relayout(wm_ctx *ctx) {
  const SB P = make_sb(ctx);
  std::vector<int> flag(P.ny), list;
  WM_HIP(hipMemcpyAsync(flag.data(), ctx->rowflag, sizeof(int) * P.ny, hipMemcpyDeviceToHost,
                        ctx->stream));
  hipLaunchKernelGGL(k_fill_i32, dim3(wm_cdiv(P.ny, 256)), dim3(256), 0, ctx->stream, ctx->rowflag,
                     (size_t)P.ny, 0);
  WM_HIP(hipStreamSynchronize(ctx->stream));
  for (int jl = 0; jl < P.ny; ++jl)
    if (flag[jl]) list.push_back(jl);
  const Mig G = wm_make_mig(ctx, ctx->mv_last);
  for (size_t r0 = 0; r0 < list.size(); r0 += ctx->relay_rows) {
    const int nb = (int)std::min<size_t>(ctx->relay_rows, list.size() - r0);
    WM_HIP(hipMemcpyAsync(ctx->relay_list, list.data() + r0, sizeof(int) * nb,
                          hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_relayout_plan, dim3(nb * P.nsp), dim3(SW), 0, ctx->stream, P,
                       ctx->relay_list, ctx->pcnt, ctx->bstart, ctx->rtmp_bstart, ctx->rtmp_blk,
                       ctx->derr);
    const size_t nthr = (size_t)nb * P.nsp * P.nx * DL;
    hipLaunchKernelGGL(k_relayout_move, dim3((unsigned)std::min<size_t>((nthr + 255) / 256, 65536)),
                       dim3(256), 0, ctx->stream, G, P, ctx->relay_list, ctx->rtmp_bstart,
                       ctx->rtmp_blk, ctx->mv[ctx->mv_last], ctx->bins, ctx->id_bins, ctx->rtmp,
                       ctx->rtmp_id, ctx->rowpend, nb, ctx->derr);
    hipLaunchKernelGGL(k_relayout_copy, dim3(std::max(1, std::min(256, wm_cdiv(P.nx, 4))), nb * P.nsp),
                       dim3(256), 0, ctx->stream, P, ctx->relay_list, ctx->rtmp_bstart,
                       ctx->rtmp_blk, ctx->rtmp, ctx->rtmp_id, ctx->bins, ctx->id_bins,
                       ctx->bin_cnt, ctx->bstart, ctx->rowpend);
    WM_HIP(hipGetLastError());
    WM_HIP(hipStreamSynchronize(ctx->stream));       // the list is host memory of this scope
  }
  ctx->relayouts += (long long)list.size();
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
    // the FP32 form keeps its particles in cell bins (wm_k_stage_to_bins)
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
