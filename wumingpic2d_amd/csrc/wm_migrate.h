// wm_migrate.h -- the FP32 form's migration of one cell bin: the particles
// that left their cells in a push join their destination bins.  Used by the
// settle kernel (k_migrate) and by the next push's prologue.
#pragma once
#include "wm_internal.h"

namespace wm_mig {

struct Mig {                    // bins and mover regions of one migration
  int nx, ny, nxr, CB;          // nxr: regions per row (cells rounded up to the push chunk)
  int r0, nr;                   // the migrated rows (every species; settle kernel)
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

// The migration of bin d (cell i, row jl, species sp), by a group of 8
// consecutive lanes (lane k of the group passes k) that alone writes d: no
// lane reads what another group writes.  Lane k takes the arrivals from the
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
// Returns d's new count to lane 0 of the group (-1 to the others).
__device__ inline int migrate_bin(const Mig &G, const Movers &M, const Bins32 &A, int64_t *ids,
                                  int *derr, size_t d, int i, int jl, int sp, int k) {
  const int dir = k < 4 ? k : k + 1;
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
  const int n1 = n - m, nnew = n1 + total;
  const size_t base = d * G.CB;
  // Lane 0's compaction (fewer arrivals than holes): the particles above the
  // new count that are not holes fill the holes left below it, in order.
  // Its sources lie at or above the new count and the arrivals' holes below
  // it, so the first NMV moves load together with the arrivals' records
  // (the rest follow one by one).
  constexpr int NMV = 4, NAR = 2;
  const bool compact = k == 0 && total < m;
  int hb = total, q = total, h = 0, from_slot = nnew;
  if (compact) {
    while (hb < m && hole_slot(M, r0, ol, hb) < nnew) ++hb;   // holes below the new count
    h = hb;
  }
  int mdst[NMV], msrc[NMV];
  bool mok[NMV];
  float4 mq[NMV];
  float muz[NMV];
  int64_t mid[NMV];
#pragma unroll
  for (int t = 0; t < NMV; ++t) {
    mok[t] = false;
    mdst[t] = msrc[t] = 0;
    if (compact) {
      while (from_slot < n && h < m && hole_slot(M, r0, ol, h) == from_slot) {   // holes above
        ++h;
        ++from_slot;
      }
      if (from_slot < n && q < hb) {
        mdst[t] = hole_slot(M, r0, ol, q);
        msrc[t] = from_slot;
        mok[t] = true;
        ++q;
        ++from_slot;
      }
    }
    if (mok[t]) {
      mq[t] = A.q[base + msrc[t]];
      muz[t] = A.uz[base + msrc[t]];
      mid[t] = ids[base + msrc[t]];
    }
  }
  // Arrivals from direction k, in the source's list order: the first NAR
  // records load together, the rest one by one.
  int a = inc - c, left = c, j = 0;
  int ms = 0;
  size_t rs = 0;
  if (c > 0) {
    ms = (int)(fh.z & 0xFFu);
    rs = bin_records(G, si, sjl, fh.z);
  }
  auto dir_of = [&](int jj) {
    return jj < WM_SUM_MV ? sum_mv(fl.x, fl.y, fl.z, fl.w, jj) >> WM_SLOT_BITS
                          : M.rec[rs + jj].sd >> 24;
  };
  auto slot_of = [&](int aa) { return aa < m ? hole_slot(M, r0, ol, aa) : n + (aa - m); };
  auto put = [&](const MoverRec &mr, int slot) {
    if (slot >= G.CB) {
      atomicOr(derr, DERR_NP_OVER);
      return;
    }
    const size_t p = base + slot;
    A.q[p] = make_float4(mr.ox1, mr.oy1, mr.ux, mr.uy);
    A.uz[p] = mr.uz;
    ids[p] = mr.id;
  };
  MoverRec mr[NAR];
  int aslot[NAR];
  bool aok[NAR];
#pragma unroll
  for (int t = 0; t < NAR; ++t) {
    aok[t] = false;
    aslot[t] = 0;
    while (left > 0 && j < ms && dir_of(j) != k) ++j;
    if (left > 0 && j < ms) {
      mr[t] = M.rec[rs + j];
      aslot[t] = slot_of(a);
      aok[t] = true;
      --left;
      ++a;
      ++j;
    }
  }
#pragma unroll
  for (int t = 0; t < NAR; ++t)
    if (aok[t]) put(mr[t], aslot[t]);
  for (; left > 0 && j < ms; ++j) {
    if (dir_of(j) != k) continue;
    put(M.rec[rs + j], slot_of(a));
    --left;
    ++a;
  }
#pragma unroll
  for (int t = 0; t < NMV; ++t) {
    if (mok[t]) {
      const size_t dst = base + mdst[t];
      A.q[dst] = mq[t];
      A.uz[dst] = muz[t];
      ids[dst] = mid[t];
    }
  }
  if (k != 0) return -1;
  if (compact) {
    for (; from_slot < n && q < hb; ++from_slot) {
      if (h < m && hole_slot(M, r0, ol, h) == from_slot) {     // a hole above the new count
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
  return cn;
}

}  // namespace wm_mig
