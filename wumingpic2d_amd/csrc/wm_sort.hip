// wm_sort.hip -- stable per-row bucket sort by x-cell (sort.f90:36-82).
//
// One workgroup of NWS wavefronts per (species, row) segment.  The cell
// histogram and the running write offsets live in LDS (nx counters): the
// wavefronts count the segment's keys with LDS atomics (tile t of 64
// particles by wavefront t % NWS, BT tiles in flight), the workgroup scans
// the counts into cumcnt, and the scatter then walks the segment in rounds
// of NWS consecutive tiles, wavefront w taking tile w of the round.  Inside a
// tile, lanes with equal keys are grouped by a ballot loop over the distinct
// keys (sorted input has one to three per tile), which gives each lane its
// rank among equal keys below it; each wavefront lists its tile's (key,
// count) pairs in LDS, and a lane adds the counts of its key in the round's
// lower tiles, so the destination of every particle is
//   start[key] + (equal keys of earlier rounds) + (of lower tiles) + rank,
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

constexpr int SW = 64;     // wavefront
constexpr int NWS = 8;     // wavefronts per segment
constexpr int STH = NWS * SW;
constexpr int BT = 4;      // tiles in flight per wavefront (count)

struct SP {
  int nx, nxgs, cap;
};

template <class R>
__global__ __launch_bounds__(STH) void k_sort(SP P, const int *__restrict__ np2, R src,
                                              const int64_t *__restrict__ ids, R dst,
                                              int64_t *__restrict__ idd,
                                              int *__restrict__ cumcnt, int *derr) {
  extern __shared__ int cnt[];             // nx counters, then running offsets
  __shared__ int lkey[NWS][SW], lcnt[NWS][SW], nl[NWS], wsum[NWS];
  using V = typename R::Val;
  const int seg = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = np2[seg];
  const int nt = (n + SW - 1) / SW;
  const size_t base = (size_t)seg * P.cap;
  int bad = 0;
  for (int i = threadIdx.x; i < P.nx; i += STH) cnt[i] = 0;
  __syncthreads();
  // count (sort.f90:59-62)
  for (int t0 = w; t0 < nt; t0 += NWS * BT) {
    int key[BT];
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = (t0 + b * NWS) * SW + lane;
      key[b] = k < n ? src.key(src.load(base + k), P.nxgs) : -1;
    }
#pragma unroll
    for (int b = 0; b < BT; ++b) {
      const int k = (t0 + b * NWS) * SW + lane;
      bool act = k < n;
      if (act && (key[b] < 0 || key[b] >= P.nx)) {
        bad = 1;
        act = false;
      }
      const WaveMatch m = wave_match(key[b], act);
      if (act && m.leader) atomicAdd(&cnt[key[b]], m.count);
    }
  }
  __syncthreads();
  // exclusive prefix -> cumcnt(nxs:nxe+1) (sort.f90:64-69); cnt becomes the
  // running write offset of each cell.  Thread th owns a run of per cells.
  {
    const int per = (P.nx + STH - 1) / STH;
    const int lo = min(P.nx, (int)threadIdx.x * per), hi = min(P.nx, lo + per);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    int incl = s;
#pragma unroll
    for (int off = 1; off < SW; off <<= 1) {
      const int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    if (lane == SW - 1) wsum[w] = incl;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < NWS; ++v) {
      pre += v < w ? wsum[v] : 0;
      tot += wsum[v];
    }
    int run = pre + incl - s;
    int *cc = cumcnt + (size_t)seg * (P.nx + 1);
    for (int i = lo; i < hi; ++i) {
      const int c = cnt[i];
      cc[i] = run;
      cnt[i] = run;
      run += c;
    }
    if (threadIdx.x == 0) cc[P.nx] = tot;
  }
  __syncthreads();
  // stable scatter (sort.f90:71-75): rounds of NWS tiles, the next round's
  // tile loading while this one is placed
  V v{}, nv{};
  int64_t vid = 0, nvid = 0;
  int key = -1, nkey = -1;
  {
    const int k = w * SW + lane;
    if (k < n) {
      v = src.load(base + k);
      vid = ids[base + k];
      key = src.key(v, P.nxgs);
    }
  }
  for (int r0 = 0; r0 < nt; r0 += NWS) {
    {
      const int k = (r0 + NWS + w) * SW + lane;
      nkey = -1;
      if (k < n) {
        nv = src.load(base + k);
        nvid = ids[base + k];
        nkey = src.key(nv, P.nxgs);
      }
    }
    const bool act = key >= 0 && key < P.nx;
    const WaveMatch m = wave_match(key, act);
    const unsigned long long lm = __ballot(act && m.leader);
    if (act && m.leader) {
      const int li = __popcll(lm & ((1ULL << lane) - 1ULL));
      lkey[w][li] = key;
      lcnt[w][li] = m.count;
    }
    if (lane == 0) nl[w] = __popcll(lm);
    __syncthreads();
    if (act) {
      int prior = 0;                       // equal keys in the round's lower tiles
      for (int u = 0; u < w; ++u) {
        const int nu = nl[u];
        for (int i = 0; i < nu; ++i) prior += lkey[u][i] == key ? lcnt[u][i] : 0;
      }
      const size_t d = base + cnt[key] + prior + m.rank;
      dst.store(d, v);
      idd[d] = vid;
    }
    __syncthreads();                       // every lane has read the offsets and lists
    if (act && m.leader) atomicAdd(&cnt[key], m.count);
    v = nv;
    vid = nvid;
    key = nkey;
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

// The re-layout's lists and workspace (wm_ctx::relay_*).
struct Relay {
  int *ctl;                      // [0] listed blocks, [1] listed segments
  int *blocks, *segs, *blkflag, *segflag;
  Bins32 ws;
  int64_t *ws_id;
  int *ws_bs;                    // per segment workgroup: new layout, count prefix
  int wsz;                       // slots per block workgroup
};

// After a push: every bin's count once the push's movers are placed
// (count - movers out + arrivals from the 8 neighbours' summaries), into pc.
// A bin whose count would exceed its room lists its block of WM_RB cells for
// the re-layout; a count above WM_BIN_MAX is the reference's memory over.
// Also keeps the capacity statistics.
__global__ __launch_bounds__(256) void k_fill_check(Mig G, int *__restrict__ pc, Relay R,
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
    if (cnew > WM_BIN_MAX) atomicOr(derr, DERR_NP_OVER);
    if (over) {
      const int nblk = (G.nx + WM_RB - 1) / WM_RB;
      const int b = (int)sj * nblk + i / WM_RB;
      if (atomicOr(R.blkflag + b, 1) == 0) R.blocks[atomicAdd(R.ctl, 1)] = b;
    }
  }
  // maxima: one lane per wave, and an atomic only where it raises the value
  // (atomics on one address from every wave would serialise)
  int a = cnew, m = mout;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a = max(a, __shfl_xor(a, o, 64));
    m = max(m, __shfl_xor(m, o, 64));
  }
  const unsigned long long fb = __ballot(over);
  if ((threadIdx.x & 63) == 0) {
    if (a > stats[0]) atomicMax(stats + 0, a);
    if (m > stats[1]) atomicMax(stats + 1, m);
    if (fb) atomicAdd(stats + 5, __popcll(fb));
  }
  if (d < (size_t)G.ny) {
    if (mfill[d] > stats[3]) atomicMax(stats + 3, mfill[d]);
    if (16 * sfill[d] > stats[4]) atomicMax(stats + 4, 16 * sfill[d]);   // blocks of 16
  }
}

// Copies the n(j) particles of bins j < nb (old first slot ob(j), new first
// slot nb_(j), count prefix pre(j)) of segment seg through the workspace W:
// first all of them out, flat in prefix order, then all back at their new
// places.  Bins move verbatim, holes included: slot k of a bin stays slot k.
// Every thread of the workgroup calls it.
template <class OB, class NB, class PRE>
__device__ void move_bins(const SB &P, int seg, int nb, OB ob, NB nb_, PRE pre, Bins32 A,
                          int64_t *ids, Bins32 W, int64_t *wid) {
  const int T = pre(nb);
  const size_t s0 = seg_at(P, seg);
  const int stride = blockDim.x;
  // flat index f -> bin j: f grows by the block size, j follows it
  int j = 0;
  for (int f = threadIdx.x; f < T; f += stride) {
    while (f >= pre(j + 1)) ++j;
    const size_t src = s0 + ob(j) + (f - pre(j));
    W.q[f] = A.q[src];
    W.uz[f] = A.uz[src];
    if (ids) wid[f] = ids[src];
  }
  __syncthreads();
  j = 0;
  for (int f = threadIdx.x; f < T; f += stride) {
    while (f >= pre(j + 1)) ++j;
    const size_t dst = s0 + nb_(j) + (f - pre(j));
    A.q[dst] = W.q[f];
    A.uz[dst] = W.uz[f];
    if (ids) ids[dst] = wid[f];
  }
  __syncthreads();
}

// Step 1: each listed block (WM_RB cells; a workgroup of 256 threads per
// block at a time) is laid out anew inside the slots its bins span now, by
// wm_bins' policy on the counts max(count now, count once the movers are
// placed): wave 0 plans, the workgroup moves the bins.  A block whose span
// cannot hold its needs with 1/16 of it to spare, or whose particles exceed
// the workspace, lists its segment for step 2.
constexpr int RW = 256;   // threads of a re-layout workgroup
__global__ __launch_bounds__(RW) void k_relayout_blocks(SB P, Relay R, const int *__restrict__ cnt,
                                                        const int *__restrict__ pc,
                                                        int *__restrict__ bstart, Bins32 A,
                                                        int64_t *__restrict__ ids) {
  __shared__ int s_old[WM_RB], s_new[WM_RB], s_pre[WM_RB + 1];
  __shared__ int s_go;
  const int nblk = (P.nx + WM_RB - 1) / WM_RB;
  const int nlist = R.ctl[0];
  Bins32 W{R.ws.q + (size_t)blockIdx.x * R.wsz, R.ws.uz + (size_t)blockIdx.x * R.wsz};
  int64_t *wid = R.ws_id ? R.ws_id + (size_t)blockIdx.x * R.wsz : nullptr;
  for (int e = blockIdx.x; e < nlist; e += gridDim.x) {
    const int b = R.blocks[e];
    const int seg = b / nblk, i0 = (b % nblk) * WM_RB, i1 = min(P.nx, i0 + WM_RB), nb = i1 - i0;
    int *bs = bstart + (size_t)seg * (P.nx + 1);
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x, i = i0 + lane;
      const bool in = lane < nb;
      const int n = in ? cnt[(size_t)seg * P.nx + i] : 0;
      const int c = in ? max(n, pc[(size_t)seg * P.nx + i]) : 0;
      const int lo = bs[i0], hi = bs[i1];
      long long C = c;
      int T = n;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        C += __shfl_xor(C, o, 64);
        T += __shfl_xor(T, o, 64);
      }
      const int mean = (int)(C / nb);
      long long S1 = in ? wm_bins::need_of(c, mean) : 0, S0 = in ? wm_bins::need_of(c, 0) : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        S1 += __shfl_xor(S1, o, 64);
        S0 += __shfl_xor(S0, o, 64);
      }
      // a block left with less than 1/16 of its span free would fill again
      // within steps: its whole segment is laid out anew instead, spreading
      // the row's free slots over every bin
      wm_bins::Policy pol;
      const bool ok = wm_bins::make_policy(C, S1, S0, nb, hi - lo, pol) && T <= R.wsz &&
                      16 * pol.units * WM_BIN_ALIGN >= hi - lo;
      const int nd = in ? wm_bins::need_of(c, pol.floor) : 0;
      int incl = nd, ninc = n;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64), w = __shfl_up(ninc, o, 64);
        if (lane >= o) {
          incl += v;
          ninc += w;
        }
      }
      if (in) {
        s_old[lane] = bs[i];
        s_new[lane] = lo + wm_bins::start_of(pol, incl - nd);
        s_pre[lane + 1] = ninc;
      }
      if (lane == 0) {
        s_pre[0] = 0;
        s_go = ok;
        R.blkflag[b] = 0;
        if (!ok && atomicOr(R.segflag + seg, 1) == 0) R.segs[atomicAdd(R.ctl + 1, 1)] = seg;
      }
    }
    __syncthreads();
    if (s_go) {
      move_bins(
          P, seg, nb, [&](int j) { return s_old[j]; }, [&](int j) { return s_new[j]; },
          [&](int j) { return s_pre[j]; }, A, ids, W, wid);
      if (threadIdx.x < nb) bs[i0 + threadIdx.x] = s_new[threadIdx.x];
    }
    __syncthreads();
  }
}

// Step 2: each listed segment (a workgroup at a time) is laid out anew as a
// whole (wm_bins::wave_layout on the same counts) and its bins moved through
// a workspace of a segment's slots.  A segment whose counts do not fit is
// the reference's memory over.
__global__ __launch_bounds__(RW) void k_relayout_segs(SB P, Relay R, const int *__restrict__ cnt,
                                                      const int *__restrict__ pc,
                                                      int *__restrict__ bstart, Bins32 A,
                                                      int64_t *__restrict__ ids, int *derr) {
  __shared__ int s_ok;
  const int nlist = R.ctl[1];
  const size_t wsz = (size_t)P.cap;
  Bins32 W{R.ws.q + (size_t)blockIdx.x * wsz, R.ws.uz + (size_t)blockIdx.x * wsz};
  int64_t *wid = R.ws_id ? R.ws_id + (size_t)blockIdx.x * wsz : nullptr;
  int *tb = R.ws_bs + (size_t)blockIdx.x * 2 * (P.nx + 1), *pre = tb + P.nx + 1;
  for (int e = blockIdx.x; e < nlist; e += gridDim.x) {
    const int seg = R.segs[e];
    const int *cn = cnt + (size_t)seg * P.nx, *pn = pc + (size_t)seg * P.nx;
    int *bs = bstart + (size_t)seg * (P.nx + 1);
    if (threadIdx.x < 64) {
      const bool ok = wm_bins::wave_layout([&](int i) { return max(cn[i], pn[i]); }, P.nx, P.cap, tb);
      // count prefix for the flat copy, 64 bins at a time
      int run = 0;
      for (int i0 = 0; i0 < P.nx; i0 += 64) {
        const int i = i0 + threadIdx.x;
        const int n = i < P.nx ? cn[i] : 0;
        int incl = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int v = __shfl_up(incl, o, 64);
          if ((int)threadIdx.x >= o) incl += v;
        }
        if (i < P.nx) pre[i + 1] = run + incl;
        run += __shfl(incl, 63, 64);
      }
      if (threadIdx.x == 0) {
        pre[0] = 0;
        s_ok = ok;
        R.segflag[seg] = 0;
        if (!ok) atomicOr(derr, DERR_NP_OVER);
      }
    }
    __syncthreads();
    if (s_ok) {
      move_bins(
          P, seg, P.nx, [&](int j) { return bs[j]; }, [&](int j) { return tb[j]; },
          [&](int j) { return pre[j]; }, A, ids, W, wid);
      for (int i = threadIdx.x; i <= P.nx; i += blockDim.x) bs[i] = tb[i];
    }
    __syncthreads();
  }
}

// Step 3: the lists' totals into the statistics, the lists emptied.
__global__ void k_relayout_done(Relay R, int *__restrict__ stats) {
  if (threadIdx.x == 0) {
    stats[6] += R.ctl[0];
    stats[7] += R.ctl[1];
    R.ctl[0] = 0;
    R.ctl[1] = 0;
  }
}

// The settle kernel: the migration of every bin of rows r0 .. r0 + nr - 1 from
// the last push's summaries and records, when no push follows to do it in
// its prologue (state download, energy sums); rows laid out anew since that
// push hold their movers already.
constexpr int MIG_HC = 120;   // holes cached in LDS per bin group
__global__ __launch_bounds__(256) void k_migrate(Mig G, Movers M, Bins32 A,
                                                 int64_t *__restrict__ ids,
                                                 const int *__restrict__ rowpend, int *derr) {
  __shared__ int hcs[256 / DL * MIG_HC];
  int *hc = hcs + (threadIdx.x / DL) * MIG_HC;
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
    (void)migrate_bin(G, M, A, ids, derr, d, i, jl, sp, (int)(t % DL), base, capd, hc, MIG_HC);
  }
}

// The IDs of the active particles of one segment per workgroup, summed and
// xor-ed (64-bit integer wrap: the order does not matter): bins of the FP32
// form (bs, cnt) or rows of the FP64 form (np2).
__global__ __launch_bounds__(256) void k_id_checksum(const int64_t *__restrict__ ids, int nx,
                                                     int cap, const int *__restrict__ bstart,
                                                     const int *__restrict__ cnt,
                                                     const int *__restrict__ np2,
                                                     unsigned long long *out) {
  const int seg = blockIdx.x;
  unsigned long long sum = 0, x = 0;
  const int64_t *row = ids + (size_t)seg * cap;
  if (bstart) {
    const int *bs = bstart + (size_t)seg * (nx + 1);
    const int *c = cnt + (size_t)seg * nx;
    for (int i = threadIdx.x >> 6; i < nx; i += blockDim.x >> 6)
      for (int k = threadIdx.x & 63; k < c[i]; k += 64) {
        const unsigned long long v = (unsigned long long)row[bs[i] + k];
        sum += v;
        x ^= v;
      }
  } else {
    for (int k = threadIdx.x; k < np2[seg]; k += blockDim.x) {
      const unsigned long long v = (unsigned long long)row[k];
      sum += v;
      x ^= v;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sum += (unsigned long long)__shfl_xor((long long)sum, o, 64);
    x ^= (unsigned long long)__shfl_xor((long long)x, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(out, sum);
    atomicXor(out + 1, x);
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

namespace {
Relay make_relay(const wm_ctx *ctx) {
  Relay R;
  R.ctl = ctx->relay_ctl;
  R.blocks = ctx->relay_blocks;
  R.segs = ctx->relay_segs;
  R.blkflag = ctx->blkflag;
  R.segflag = ctx->segflag;
  R.ws = ctx->rws;
  R.ws_id = ctx->rws_id;
  R.ws_bs = ctx->rws_bs;
  R.wsz = ctx->relay_wsz;
  return R;
}
}  // namespace

int wm_k_fill_check(wm_ctx *ctx) {
  const Mig G = wm_make_mig(ctx, ctx->mv_last);
  const size_t nthr = std::max(G.nbin, (size_t)ctx->g.ny);
  hipLaunchKernelGGL(k_fill_check, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, ctx->stream,
                     G, ctx->pcnt, make_relay(ctx), ctx->mfill[ctx->mv_last], ctx->sfill,
                     ctx->fstat, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// The blocks the fill check listed, then the segments they escalated, laid
// out anew on the device (no host round trip: the grids loop over the lists,
// whose lengths only the device knows).
int wm_relayout(wm_ctx *ctx) {
  const SB P = make_sb(ctx);
  const Relay R = make_relay(ctx);
  hipLaunchKernelGGL(k_relayout_blocks, dim3(ctx->relay_nwb), dim3(RW), 0, ctx->stream, P, R,
                     ctx->bin_cnt, ctx->pcnt, ctx->bstart, ctx->bins, ctx->id_bins);
  hipLaunchKernelGGL(k_relayout_segs, dim3(ctx->relay_nws), dim3(RW), 0, ctx->stream, P, R,
                     ctx->bin_cnt, ctx->pcnt, ctx->bstart, ctx->bins, ctx->id_bins, ctx->derr);
  hipLaunchKernelGGL(k_relayout_done, dim3(1), dim3(64), 0, ctx->stream, R, ctx->fstat);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_id_checksum(wm_ctx *ctx, unsigned long long *out_dev) {
  const bool f32 = ctx->form == WM_FORM_FP32;
  const int64_t *ids = f32 ? ctx->id_bins : ctx->id[ctx->idcur];
  if (!ids) {
    wm_set_error(ctx, WM_ERR_ARG, "ID checksum: the particles carry no IDs (ndim = 5)");
    return WM_ERR_ARG;
  }
  hipLaunchKernelGGL(k_id_checksum, dim3(ctx->g.nsp * ctx->g.ny), dim3(256), 0, ctx->stream, ids,
                     ctx->g.nx, ctx->cap, f32 ? ctx->bstart : nullptr,
                     f32 ? ctx->bin_cnt : nullptr, ctx->np2, out_dev);
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
    // the FP32 form keeps its particles in cell bins (wm_k_stage_to_bins)
    wm_set_error(ctx, WM_ERR_ARG, "sort__bucket works on the FP64 particle form");
    return WM_ERR_ARG;
  }
  {
    hipLaunchKernelGGL(k_sort<Rec64>, dim3(nseg), dim3(STH), lds, ctx->stream, P, ctx->np2,
                       ctx->b64(src), ctx->id[ctx->idcur], ctx->b64(dst), ctx->id[1 - ctx->idcur],
                       ctx->cumcnt, ctx->derr);
  }
  WM_HIP(hipGetLastError());
  ctx->idcur = 1 - ctx->idcur;
  return WM_OK;
}
