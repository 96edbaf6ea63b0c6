// wm_records.h -- the two particle record forms kept in HBM, as
// structure-of-arrays banks, and the per-record operations the migration and
// sort kernels share.
//
//  Rec64 (FP64 form, the reference's arithmetic): absolute x, y and
//        u = gamma v in FP64 -- 40 B per particle (+ 8 B ID).
//  Rec32 (FP32 form): cell index ci (int32, x-cell relative to nxgs), the
//        in-cell offsets ox, oy in [0, 1), u in FP32 -- 24 B per particle
//        (+ 8 B ID).  In the cell bins of the resident state the cell index
//        is the bin's and ci is not stored (20 B).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct Rec64 {
  double *x, *y, *ux, *uy, *uz;
  struct Val {
    double x, y, ux, uy, uz;
  };
  __device__ inline Val load(size_t p) const { return Val{x[p], y[p], ux[p], uy[p], uz[p]}; }
  __device__ inline void store(size_t p, const Val &v) const {
    x[p] = v.x;
    y[p] = v.y;
    ux[p] = v.ux;
    uy[p] = v.uy;
    uz[p] = v.uz;
  }
  // x-cell relative to nxgs, as sort.f90:60 bins it (int(x), delx = 1)
  __device__ inline int key(const Val &v, int nxgs) const { return (int)v.x - nxgs; }
};

struct Rec32 {
  int *ci;
  float *ox, *oy, *ux, *uy, *uz;
  struct Val {
    int ci;
    float ox, oy, ux, uy, uz;
  };
  __device__ inline Val load(size_t p) const { return Val{ci[p], ox[p], oy[p], ux[p], uy[p], uz[p]}; }
  __device__ inline void store(size_t p, const Val &v) const {
    ci[p] = v.ci;
    ox[p] = v.ox;
    oy[p] = v.oy;
    ux[p] = v.ux;
    uy[p] = v.uy;
    uz[p] = v.uz;
  }
  __device__ inline int key(const Val &v, int) const { return v.ci; }
};

// Largest float below 1: offsets are kept in [0, 1) after a cell change.
__device__ inline float wm_below_one(float v) { return v < 1.0f ? v : 0x1.fffffep-1f; }

// FP32 form, cell bins: per slot a float4 (ox, oy, ux, uy) and uz (the IDs
// apart).  A particle placed by the migration touches three arrays.
struct Bins32 {
  float4 *q = nullptr;          // offsets in the cell, in-plane momenta
  float *uz = nullptr;
};

// Fixed-order sum over a 64-lane wave: a butterfly on lane xor 32, 16, ...,
// 1.  Partners add the same two values (IEEE addition commutes), so every
// lane ends with the bit-identical total.
__device__ inline double wm_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Periodic index a mod n for a within a few periods of [0, n): no division
// on the common in-range path.
__device__ inline int wm_wrap(int a, int n) {
  if ((unsigned)a < (unsigned)n) return a;
  if (a < 0 && a + n >= 0) return a + n;
  if (a >= n && a - n < n) return a - n;
  a %= n;
  return a < 0 ? a + n : a;
}
