// wm_migrate.h -- the FP32 form's migration of one cell bin: the particles
// that left their cells in a push join their destination bins.  Used by the
// settle kernel (k_migrate) and by the next push's prologue.
#pragma once
#include "wm_internal.h"

namespace wm_mig {

// Inlined into the caller, so that the structs below stay in registers.
#define WM_MIG_FN __device__ __attribute__((always_inline)) inline

struct Mig {                    // bins and mover pools of one migration
  int nx, ny, cap, mrow;        // cap: slots per row segment; mrow: records per row pool
  int r0, nr;                   // the migrated rows (every species; settle kernel)
  size_t nbin;                  // bins of those rows
  const BinSum *sum;            // per bin: the push's summary of the movers that left it
  const int *bstart;            // [seg][nx + 1] first slot of each bin in its segment
  int *cnt;
};

// First record of the movers of a bin of row jl in the row's pool (first:
// the bin summary's index).
WM_MIG_FN size_t bin_records(const Mig &G, int jl, uint32_t first) {
  return (size_t)jl * G.mrow + first;
}
// First slot and capacity of bin i of segment seg.
WM_MIG_FN void bin_span(const Mig &G, size_t seg, int i, size_t &base, int &capd) {
  const int *b = G.bstart + seg * (size_t)(G.nx + 1) + i;
  const int b0 = b[0];
  base = seg * (size_t)G.cap + b0;
  capd = b[1] - b0;
}

// Byte k of a summary's direction counts (words lo, hi), and entry j of its
// mover list (words w0..w3).  The words are values, not a dynamically indexed
// array, which the compiler would place in memory.
WM_MIG_FN int sum_cnt(uint32_t lo, uint32_t hi, int k) {
  const uint32_t v = k < 4 ? lo : hi;
  return (int)((v >> (8 * (k & 3))) & 0xFFu);
}
WM_MIG_FN int sum_mv(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int j) {
  const uint32_t v = j < 4 ? (j < 2 ? w0 : w1) : (j < 6 ? w2 : w3);
  return (int)((v >> (16 * (j & 1))) & 0xFFFFu);
}
// Bin d's holes (its movers' slots, ascending): the first WM_SUM_MV from its
// summary (words l), the next mc from the group's LDS cache hc (read from
// the records once, by the whole group), the rest from the records.
struct Holes {
  const Movers *M;
  size_t r0;                    // d's first record in its row's pool
  uint4 l;
  const int *hc;
  int mc;
};
WM_MIG_FN int hole_slot(const Holes &H, int h) {
  if (h < WM_SUM_MV) return sum_mv(H.l.x, H.l.y, H.l.z, H.l.w, h) & ((1 << WM_SLOT_BITS) - 1);
  if (h - WM_SUM_MV < H.mc) return H.hc[h - WM_SUM_MV];
  return H.M->rec[H.r0 + h].sd & 0xFFFFFF;
}

// The compaction's running state: next hole to fill (q), end of the holes
// below the new count (hb), next hole above it (h), next source slot (from).
struct MoveState {
  int q, hb, h, from;
};
struct Move {                    // one compaction move: particle at src -> slot dst
  int dst, src;
  bool ok;
  float qx, qy, qz, qw, uz;      // components, not a float4: the vector type's copy
  int64_t id;                    // would keep the struct in memory
};
// Plans the next move (skipping holes above the new count) and issues its
// source's loads.
WM_MIG_FN void plan_move(Move &v, MoveState &st, bool compact, const Holes &H, int m, int n,
                         size_t base, const Bins32 &A, const int64_t *ids) {
  v.ok = false;
  v.dst = v.src = 0;
  if (compact) {
    while (st.from < n && st.h < m && hole_slot(H, st.h) == st.from) {
      ++st.h;
      ++st.from;
    }
    if (st.from < n && st.q < st.hb) {
      v.dst = hole_slot(H, st.q);
      v.src = st.from;
      v.ok = true;
      ++st.q;
      ++st.from;
    }
  }
  if (v.ok) {
    const float4 t = A.q[base + v.src];
    v.qx = t.x;
    v.qy = t.y;
    v.qz = t.z;
    v.qw = t.w;
    v.uz = A.uz[base + v.src];
    v.id = ids ? ids[base + v.src] : 0;
  }
}
WM_MIG_FN void store_move(const Move &v, const Bins32 &A, int64_t *ids, size_t base) {
  if (!v.ok) return;
  A.q[base + v.dst] = make_float4(v.qx, v.qy, v.qz, v.qw);
  A.uz[base + v.dst] = v.uz;
  if (ids) ids[base + v.dst] = v.id;
}

// One arrival: the record and the slot it takes.
struct Arrival {
  MoverRec rec;
  int slot;
  bool ok;
};
WM_MIG_FN void put_arrival(const Arrival &w, int capd, const Bins32 &A, int64_t *ids, int *derr,
                           size_t base) {
  if (!w.ok) return;
  if (w.slot >= capd) {
    atomicOr(derr, DERR_NP_OVER);
    return;
  }
  const size_t p = base + w.slot;
  A.q[p] = make_float4(w.rec.ox1, w.rec.oy1, w.rec.ux, w.rec.uy);
  A.uz[p] = w.rec.uz;
  if (ids) ids[p] = w.rec.id;
}

// Source bin of direction index k (directions 0, 1, 2, 3, 5, 6, 7, 8) for
// the bin in cell (i, jl) of species sp.
WM_MIG_FN size_t source_bin(const Mig &G, int i, int jl, int sp, int k, int &si, int &sjl) {
  const int dir = k < 4 ? k : k + 1;
  si = wm_wrap(i - (dir % 3 - 1), G.nx);
  sjl = wm_wrap(jl - (dir / 3 - 1), G.ny);
  return ((size_t)sp * G.ny + sjl) * G.nx + si;
}

// The migration of bin d (cell i, row jl, species sp; first slot base,
// capd slots), by a group of 8 consecutive lanes (lane k of the group
// passes k) that alone writes d: no lane reads what another group writes.
// Lane k looks up the arrivals from the source bin in direction index k:
// their count (the source's summary) and where they are (a source with at
// most WM_SUM_MV movers lists them in its summary; one with more lists its
// records of one direction together through its perm index, in list order).  Arrival index = the earlier directions' count (a prefix over the
// group's lanes) + the rank within direction k, and the group's lanes take
// arrival indices k, k + 8, ..., whatever their direction.  d held n
// particles and lost m (its movers' slots, ascending, are its holes): the
// a-th arrival takes the a-th hole while there are holes, then slot
// n + (a - m).  With fewer arrivals than holes, lane 0 fills the holes left
// below the new count, in order, with the particles above it.  The
// summaries of d and of the source (32 bytes each) give the counts, the
// record offsets and the first WM_SUM_MV holes; only a bin with more movers
// reads the rest of its holes from its records.
constexpr int DL = 8;
// hc: the group's LDS cache of hcn holes past the summary's (nullptr: none).
// Returns d's new count to lane 0 of the group (-1 to the others).
WM_MIG_FN int migrate_bin(const Mig &G, const Movers &M, const Bins32 &A, int64_t *ids, int *derr,
                          size_t d, int i, int jl, int sp, int k, size_t base, int capd,
                          int *hc, int hcn) {
  int si, sjl;
  const size_t src = source_bin(G, i, jl, sp, k, si, sjl);
  const uint4 *own = (const uint4 *)(G.sum + d), *from = (const uint4 *)(G.sum + src);
  uint4 oh = own[0], ol = own[1], fh = from[0], fl = from[1];
  int nd = G.cnt[d];
  // all loads in flight together (the compiler would otherwise issue some
  // only where they are first used, a round trip later)
  __asm__ volatile("" : "+v"(ol.x), "+v"(ol.y), "+v"(ol.z), "+v"(ol.w), "+v"(fh.x), "+v"(fh.y),
                   "+v"(fh.z), "+v"(fh.w), "+v"(fl.x), "+v"(fl.y), "+v"(fl.z), "+v"(fl.w),
                   "+v"(nd));
  const int n = min(nd, capd);
  const int c = sum_cnt(fh.x, fh.y, k);
  const int m = (int)oh.w;
  const size_t r0 = bin_records(G, jl, oh.z);
  int inc = c;                                   // inclusive prefix over the group
#pragma unroll
  for (int o = 1; o < DL; o <<= 1) {
    const int v = __shfl_up(inc, o, DL);
    if (k >= o) inc += v;
  }
  const int total = __shfl(inc, DL - 1, DL);
  const int n1 = n - m, nnew = n1 + total;
  // a bin with more holes than its summary lists: the group reads the next
  // ones from its records into the LDS cache, four per lane in flight
  const int mc = hc ? min(max(m - WM_SUM_MV, 0), hcn) : 0;
  for (int e0 = 0; e0 < mc; e0 += 4 * DL) {
    int hv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = e0 + u * DL + k;
      hv[u] = e < mc ? (int)(M.rec[r0 + WM_SUM_MV + e].sd & 0xFFFFFF) : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = e0 + u * DL + k;
      if (e < mc) hc[e] = hv[u];
    }
  }
  __builtin_amdgcn_wave_barrier();                 // the group reads what its lanes wrote
  const Holes H{&M, r0, ol, hc, mc};
  // Lane 0's compaction (fewer arrivals than holes): the particles above the
  // new count that are not holes fill the holes left below it, in order.
  // Its sources lie at or above the new count and the arrivals' holes below
  // it, so the first four moves load together with the arrivals' records
  // (the rest follow one by one).  Each batched move is a named variable,
  // not an array element: an indexed array would go to memory.
  const bool compact = k == 0 && total < m;
  MoveState st{total, total, 0, nnew};
  if (compact) {
    while (st.hb < m && hole_slot(H, st.hb) < nnew) ++st.hb;   // holes below the new count
    st.h = st.hb;
  }
  Move v0, v1, v2, v3;
  plan_move(v0, st, compact, H, m, n, base, A, ids);
  plan_move(v1, st, compact, H, m, n, base, A, ids);
  plan_move(v2, st, compact, H, m, n, base, A, ids);
  plan_move(v3, st, compact, H, m, n, base, A, ids);
  // direction k's records of the source: perm[fs + pre + r], r < c, index
  // records fs + perm[..]; pre = the source's movers of directions before k
  int pre = 0;
#pragma unroll
  for (int q = 0; q < DL - 1; ++q) pre += q < k ? sum_cnt(fh.x, fh.y, q) : 0;
  const size_t fs = c > 0 ? bin_records(G, sjl, fh.z) : 0;
  const unsigned fs_lo = (unsigned)fs, fs_hi = (unsigned)(fs >> 32);
  // arrival a: its direction kk (the group's inclusive prefixes ascend),
  // rank a - (the prefix before kk), and lane kk's source.  The shuffles run
  // in every lane of the group (a lane past the last arrival reads and
  // drops).
  auto arrival = [&](int a, Arrival &w) {
    w.ok = a < total;
    int kk = 0, before = 0;
#pragma unroll
    for (int q = 0; q < DL; ++q) {
      const int iq = __shfl(inc, q, DL);
      if (iq <= a) {
        kk = q + 1;
        before = iq;
      }
    }
    kk = min(kk, DL - 1);
    const size_t f = (size_t)(unsigned)__shfl((int)fs_lo, kk, DL) |
                     ((size_t)(unsigned)__shfl((int)fs_hi, kk, DL) << 32);
    const int pk = __shfl(pre, kk, DL);
    const int ms = __shfl((int)fh.w, kk, DL);
    const uint32_t l0 = __shfl(fl.x, kk, DL), l1 = __shfl(fl.y, kk, DL);
    const uint32_t l2 = __shfl(fl.z, kk, DL), l3 = __shfl(fl.w, kk, DL);
    if (!w.ok) return;
    const int r = a - before;
    int q;
    if (ms <= WM_SUM_MV) {
      // a source with few movers lists them all in its summary (direction
      // index in the top bits): the r-th of direction kk, without a load
      q = 0;
      for (int j = 0, seen = 0; j < ms; ++j)
        if (sum_mv(l0, l1, l2, l3, j) >> WM_SLOT_BITS == kk) {
          if (seen == r) q = j;
          ++seen;
        }
    } else {
      q = M.perm[f + pk + r];
    }
    w.rec = M.rec[f + q];
    w.slot = a < m ? hole_slot(H, a) : n + (a - m);
  };
  // every lane of the group runs the same number of rounds (the shuffles
  // need the whole group)
  for (int a = k; a - k < total; a += DL) {
    Arrival w;
    arrival(a, w);
    put_arrival(w, capd, A, ids, derr, base);
  }
  store_move(v0, A, ids, base);
  store_move(v1, A, ids, base);
  store_move(v2, A, ids, base);
  store_move(v3, A, ids, base);
  if (k != 0) return -1;
  // the rest of the compaction, four moves at a time (a bin that lost many
  // particles and gained few)
  while (compact && st.from < n && st.q < st.hb) {
    Move u0, u1, u2, u3;
    plan_move(u0, st, compact, H, m, n, base, A, ids);
    plan_move(u1, st, compact, H, m, n, base, A, ids);
    plan_move(u2, st, compact, H, m, n, base, A, ids);
    plan_move(u3, st, compact, H, m, n, base, A, ids);
    store_move(u0, A, ids, base);
    store_move(u1, A, ids, base);
    store_move(u2, A, ids, base);
    store_move(u3, A, ids, base);
  }
  int cn = n1 + total;
  if (cn > capd) {
    atomicOr(derr, DERR_NP_OVER);
    cn = capd;
  }
  G.cnt[d] = cn;
  return cn;
}

}  // namespace wm_mig

// The bins and the mover buffer buf of the context (wm_sort.hip).
wm_mig::Mig wm_make_mig(const wm_ctx *ctx, int buf);
