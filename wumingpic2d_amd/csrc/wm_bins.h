// wm_bins.h -- the FP32 form's bin layout of one row segment, and the
// wavefront key matching of the stable scatters that fill bins in row order.
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

// Room of a bin of c particles before the slack: at least one aligned block.
__device__ inline int bin_need(int c) { return round_up(max(c, 1), WM_BIN_ALIGN); }

// The layout of one segment of `cap` slots, computed by one wavefront (all
// 64 lanes call it) from the counts count(i), i < nx: bin i gets its need
// (count rounded up to WM_BIN_ALIGN slots, at least one block) plus a share
// of the segment's free slots in proportion to that need, in whole blocks,
// at most WM_BIN_MAX slots.  Writes the first slots bs[0..nx] (bs[nx] = slots
// laid out) and returns false when the counts do not fit: the row's
// capacity is exceeded (the reference's memory over) or one bin's count is
// above WM_BIN_MAX.
template <class F>
__device__ bool wave_layout(F count, int nx, int cap, int *bs) {
  const int lane = threadIdx.x & 63;
  const int per = (nx + 63) / 64;
  const int lo = min(nx, lane * per), hi = min(nx, lo + per);
  long long need = 0;
  int bad = 0;
  for (int i = lo; i < hi; ++i) {
    const int c = count(i);
    bad |= c > WM_BIN_MAX;
    need += bin_need(c);
  }
  long long S = need;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) S += __shfl_xor(S, o, 64);
  if (__ballot(bad) != 0ULL || S > cap) return false;
  const long long units = (cap - S) / WM_BIN_ALIGN;
  constexpr int top = WM_BIN_MAX / WM_BIN_ALIGN * WM_BIN_ALIGN;
  auto room = [&](int c) {
    const int nd = bin_need(c);
    const long long extra = units * nd / S * WM_BIN_ALIGN;
    return (int)min((long long)top, nd + extra);
  };
  int mine = 0;
  for (int i = lo; i < hi; ++i) mine += room(count(i));
  int incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  int run = incl - mine;
  for (int i = lo; i < hi; ++i) {
    bs[i] = run;
    run += room(count(i));
  }
  const int total = __shfl(incl, 63, 64);
  if (lane == 0) bs[nx] = total;
  return true;
}

}  // namespace wm_bins
