// wm_bins.h -- the FP32 form's bin layout of one row segment (or of a window
// of it), and the wavefront key matching of the stable scatters that fill
// bins in row order.
#pragma once
#include "wm_internal.h"

namespace wm_bins {

struct WaveMatch {
  int rank;     // equal keys in lower lanes
  int count;    // equal keys in the wavefront
  bool leader;  // lowest lane holding the key
};

// Groups the active lanes of a wavefront by key (a ballot loop over the
// distinct keys; inputs sorted or nearly so have few per wavefront).
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

__device__ inline int round_up(int v, int a) { return (v + a - 1) / a * a; }

// The layout policy of a window of nb bins over `span` slots: bin i gets its
// need -- its count raised to the window's mean count (floor), rounded up to
// WM_BIN_ALIGN slots, at least one block -- plus a share of the window's
// free slots in proportion to that need, in whole blocks.  Raising the small
// bins to the mean keeps a bin that is momentarily light from being laid out
// tight and overflowing as soon as its count returns to the mean; the share
// in proportion gives a dense bin, which gains particles fastest where a
// plasma pinches, headroom for as many steps as a mean one.  The shares are
// apportioned by the running sum of the needs: bin i starts at
// P + WM_BIN_ALIGN * floor(units * P / S), P the needs of the bins before it,
// so the free blocks are all handed out and a bin's share differs from its
// proportion by less than one block.  (Rounding each share down on its own
// gave nothing to every bin whose proportion was below one block: with 32
// particles per cell and rows of 1.4 times the mean, every bin at the mean
// count was laid out full.)  When the raised needs do not fit, the plain
// counts are used (floor 0).
struct Policy {
  int floor;        // counts below it are raised to it
  long long units;  // free blocks of WM_BIN_ALIGN slots after the needs
  long long S;      // the needs' sum
};
__device__ inline int need_of(int c, int floor) { return round_up(max(max(c, floor), 1), WM_BIN_ALIGN); }
// First slot (from the window's start) of the bin whose predecessors' needs
// sum to P.
__device__ inline int start_of(const Policy &p, long long P) {
  return (int)(P + (long long)WM_BIN_ALIGN * (p.units * P / p.S));
}
// From the window's count sum C, the sum of needs raised to the mean (S1)
// and the plain one (S0): false when even the plain needs exceed the span.
__device__ inline bool make_policy(long long C, long long S1, long long S0, int nb, long long span,
                                   Policy &p) {
  if (S1 <= span) {
    p.floor = (int)(C / nb);
    p.S = S1;
  } else {
    p.floor = 0;
    p.S = S0;
  }
  p.units = (span - p.S) / WM_BIN_ALIGN;
  return p.S <= span;
}

// The layout of one segment of `cap` slots, computed by one wavefront (all
// 64 lanes call it) from the counts count(i), i < nx, with the policy above.
// Writes the first slots bs[0..nx] (bs[nx] = cap) and returns false when the
// counts do not fit: the row's capacity is exceeded (the reference's memory
// over) or one bin's count is above WM_BIN_MAX.
template <class F>
__device__ bool wave_layout(F count, int nx, int cap, int *bs) {
  const int lane = threadIdx.x & 63;
  const int per = (nx + 63) / 64;
  const int lo = min(nx, lane * per), hi = min(nx, lo + per);
  long long C = 0;
  int bad = 0;
  for (int i = lo; i < hi; ++i) {
    const int c = count(i);
    bad |= c > WM_BIN_MAX;
    C += c;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) C += __shfl_xor(C, o, 64);
  if (__ballot(bad) != 0ULL) return false;
  const int mean = (int)(C / nx);
  long long S1 = 0, S0 = 0;
  for (int i = lo; i < hi; ++i) {
    const int c = count(i);
    S1 += need_of(c, mean);
    S0 += need_of(c, 0);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    S1 += __shfl_xor(S1, o, 64);
    S0 += __shfl_xor(S0, o, 64);
  }
  Policy p;
  if (!make_policy(C, S1, S0, nx, cap, p)) return false;
  long long mine = 0;
  for (int i = lo; i < hi; ++i) mine += need_of(count(i), p.floor);
  long long incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  long long run = incl - mine;                  // needs of the bins before lo
  for (int i = lo; i < hi; ++i) {
    bs[i] = start_of(p, run);
    run += need_of(count(i), p.floor);
  }
  if (lane == 0) bs[nx] = cap;     // the slack after the last block joins the last bin
  return true;
}

}  // namespace wm_bins
