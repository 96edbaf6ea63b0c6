// wm_particle32.hip -- Buneman-Boris push and Esirkepov deposit in the FP32
// form (int32 cell index + FP32 in-cell offsets + FP32 momenta).
//
// Same decomposition as the FP64 kernel: one workgroup per (row, chunk of
// CHUNK cells), a group of GROUP lanes per cell striding through the cell's
// particles in windows of GROUP consecutive particles.  The push and
// the per-lane deposit run in FP32; the group's stencils are summed in FP32
// and then converted, so the workgroup tile, the cross-tile gather and the
// field solve accumulate in FP64.
//
// Positions never leave the cell frame: the push moves the offsets
// (ox, oy) by u dt / gamma; ox is rebased on the new cell (the x boundary is
// an exact integer wrap of the cell index), oy stays relative to the old row
// until the row migration rebases it.  The shape functions take the offsets
// directly (dh = offset - 1/2), which keeps their accuracy independent of the
// grid size.
#include "wm_internal.h"

namespace {

constexpr int CHUNK = 16;             // cells per workgroup
constexpr int GROUP = 16;             // lanes per cell
constexpr int TW = CHUNK + 2;         // push tile width (halo 1)
constexpr int DW = CHUNK + 4;         // deposit tile width (halo 2)
constexpr int TILE = 5 * DW * 3;      // doubles per deposit tile
constexpr int NACC = 65;

struct PP32 {
  int nx, ny, nxgs, nys, nsp, cap, nchunk;
  float delt, c, c2, inv_c2, qc_scale;   // qc_scale = delx/delt
  float q[WM_MAXSP], r[WM_MAXSP];
};

__device__ inline void shapef(float dh, float &sm, float &s0, float &sp) {
  sm = 0.5f * (0.5f - dh) * (0.5f - dh);
  s0 = 0.75f - dh * dh;
  sp = 0.5f * (0.5f + dh) * (0.5f + dh);
}

struct Push32 {
  float ox, oy, ux, uy, uz;   // ox unnormalised (relative to the old cell)
};

// The TSC gather of particle.f90:107-129 inside one cell is a biquadratic
// in the offsets: with w-(o) = 1/2 - o + o^2/2, w0(o) = 1/2 + o - o^2,
// w+(o) = o^2/2, sum_ab T(a,b) w_a(ox) w_b(oy) = sum_ij C_ij ox^i oy^j.
// The workgroup computes C (in FP64, rounded once) for each of its cells and
// fields; a particle then needs 8 FMAs per field.
constexpr int CFS = 56;               // floats per cell: 6 fields x 9, padded to 16 B
__device__ inline void cell_coeffs(float *cf, const float *T, int c, int k) {
  const double M[3][3] = {{0.5, -1.0, 0.5}, {0.5, 1.0, -1.0}, {0.0, 0.0, 0.5}};
  const float *Tk = T + k * 3 * TW + c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double v = 0.0;
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) v += (double)Tk[b * TW + a] * M[a][i] * M[b][j];
      cf[k * 9 + i * 3 + j] = (float)v;
    }
}

// particle.f90:96-161 in FP32 on the cell frame.
__device__ inline Push32 boris32(const float4 *cf4, float ox, float oy, float ux, float uy,
                                 float uz, const PP32 &P, float fac1, float txxx, float fac2) {
  float c[CFS];
#pragma unroll
  for (int q = 0; q < CFS / 4; ++q) {
    const float4 v = cf4[q];
    c[4 * q] = v.x; c[4 * q + 1] = v.y; c[4 * q + 2] = v.z; c[4 * q + 3] = v.w;
  }
  float f[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float *C = c + 9 * k;
    const float r0 = C[0] + oy * (C[1] + oy * C[2]);
    const float r1 = C[3] + oy * (C[4] + oy * C[5]);
    const float r2 = C[6] + oy * (C[7] + oy * C[8]);
    f[k] = r0 + ox * (r1 + ox * r2);
  }
  const float bpx = f[0], bpy = f[1], bpz = f[2];
  const float epx = f[3], epy = f[4], epz = f[5];
  float uvm1 = ux + fac1 * epx;
  float uvm2 = uy + fac1 * epy;
  float uvm3 = uz + fac1 * epz;
  // hardware reciprocal square root and reciprocal (1 ulp): the FP32 form's
  // error budget is set by FP32 storage, not by these
  const float g2 = P.c2 + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3;
  const float igam = __builtin_amdgcn_rsqf(g2);
  const float gam = g2 * igam;
  const float fac1r = fac1 * igam;
  const float fac2r = fac2 * __builtin_amdgcn_rcpf(gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
  const float uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
  const float uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
  const float uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
  uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
  uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
  uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
  Push32 o;
  o.ux = uvm1 + fac1 * epx;
  o.uy = uvm2 + fac1 * epy;
  o.uz = uvm3 + fac1 * epz;
  const float ig = __builtin_amdgcn_rsqf(1.0f + (o.ux * o.ux + o.uy * o.uy + o.uz * o.uz) * P.inv_c2);
  o.ox = ox + o.ux * P.delt * ig;
  o.oy = oy + o.uy * P.delt * ig;
  return o;
}

// Cell change and the offset rebased on the new cell, kept in [0, 1).
__device__ inline int rebase(float o, float &onew) {
  const int d = o < 0.0f ? -1 : (o >= 1.0f ? 1 : 0);
  onew = d > 0 ? o - 1.0f : (d < 0 ? wm_below_one(o + 1.0f) : o);
  return d;
}

struct Move32 {
  float s0x[5], s0y[5], dsx[5], dsy[5];
};

__device__ inline void place_s1(float a, float b, float c, int inc, float *ds) {
  ds[0] = inc < 0 ? a : 0.0f;
  ds[1] = inc < 0 ? b : (inc == 0 ? a : 0.0f);
  ds[2] = inc < 0 ? c : (inc == 0 ? b : a);
  ds[3] = inc > 0 ? b : (inc == 0 ? c : 0.0f);
  ds[4] = inc > 0 ? c : 0.0f;
}

// field.f90:224-268 on offsets: old offsets (ox0, oy0) in cell (0, 0), new
// rebased offsets (ox1, oy1) in cell (dx, dy).
__device__ inline void make_move32(Move32 &m, float ox0, float oy0, float ox1, float oy1, int dx,
                                   int dy) {
  m.s0x[0] = m.s0x[4] = m.s0y[0] = m.s0y[4] = 0.0f;
  shapef(ox0 - 0.5f, m.s0x[1], m.s0x[2], m.s0x[3]);
  shapef(oy0 - 0.5f, m.s0y[1], m.s0y[2], m.s0y[3]);
  float a, b, c;
  shapef(ox1 - 0.5f, a, b, c);
  place_s1(a, b, c, dx, m.dsx);
  shapef(oy1 - 0.5f, a, b, c);
  place_s1(a, b, c, dy, m.dsy);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    m.dsx[k] -= m.s0x[k];
    m.dsy[k] -= m.s0y[k];
  }
}

template <int C>
__device__ inline void stencil32(float *acc, const Move32 &m, float qc, float qv) {
  if (C == 0) {
    float w = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      w -= qc * m.dsx[a];
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[a * 5 + b] += w * (m.s0y[b] + 0.5f * m.dsy[b]);
    }
  } else if (C == 1) {
    float w = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      w -= qc * m.dsy[b];
#pragma unroll
      for (int a = 0; a < 5; ++a) acc[a * 4 + b] += (m.s0x[a] + 0.5f * m.dsx[a]) * w;
    }
  } else {
    const float third = 1.0f / 3.0f;
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      const float za = qv * (m.s0x[a] + 0.5f * m.dsx[a]);
      const float zb = qv * (0.5f * m.s0x[a] + third * m.dsx[a]);
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[a * 5 + b] += za * m.s0y[b] + zb * m.dsy[b];
    }
  }
}

// Stay-in-cell particle: the 27 possibly non-zero stencil values (same
// layout as the FP64 kernel's core set).
__device__ inline void core32(float *core, float ox0, float oy0, float ox1, float oy1, float qc,
                              float qv) {
  float sx[3], sy[3], dx[3], dy[3];
  shapef(ox0 - 0.5f, sx[0], sx[1], sx[2]);
  shapef(oy0 - 0.5f, sy[0], sy[1], sy[2]);
  shapef(ox1 - 0.5f, dx[0], dx[1], dx[2]);
  shapef(oy1 - 0.5f, dy[0], dy[1], dy[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dx[k] -= sx[k];
    dy[k] -= sy[k];
  }
  float X[3], Y[3], wx[3], wy[3];
  float w1 = 0.0f, w2 = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] = sx[k] + 0.5f * dx[k];
    Y[k] = sy[k] + 0.5f * dy[k];
    w1 -= qc * dx[k];
    w2 -= qc * dy[k];
    wx[k] = w1;
    wy[k] = w2;
  }
  const float third = 1.0f / 3.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float za = qv * X[a];
    const float zb = qv * (0.5f * sx[a] + third * dx[a]);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      core[a * 3 + b] += wx[a] * Y[b];
      core[9 + a * 3 + b] += X[a] * wy[b];
      core[18 + a * 3 + b] += za * sy[b] + zb * dy[b];
    }
  }
}

__device__ inline int core_slot(int c) {
  const int comp = c / 9, a = (c % 9) / 3 + 1, b = c % 3 + 1;
  if (comp == 0) return a * 5 + b;
  if (comp == 1) return 20 + a * 4 + b;
  return 40 + a * 5 + b;
}

__device__ inline double acc_at(const double *cb, int comp, int a, int b) {
  if (comp == 0) return a >= 1 ? cb[(a - 1) * 5 + b] : 0.0;
  if (comp == 1) return b >= 1 ? cb[20 + a * 4 + (b - 1)] : 0.0;
  return cb[40 + a * 5 + b];
}

// Sum over the 16 lanes of one DPP row (a group): quad butterflies, then
// rotations by 4 and 8 within the row.  Every lane ends with the total; the
// callers use lane 0's, whose association order is fixed.
__device__ inline float row_sum16(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x124, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xF, 0xF, false));
  return v;
}




constexpr int XR = 64;          // listed movers staged per round of the workgroup pass

// Bin of (species s, row jl, cell i).
__device__ inline size_t bin_of(int s, int jl, int i, int ny, int nx) {
  return ((size_t)s * ny + jl) * nx + i;
}

// One component of the stencil of a cell-crossing particle from its mover
// record R: comp 0 -> out[0..19] (Jx), 1 -> out[20..39] (Jy), 2 -> out[40..64]
// (Jz).
__device__ inline void mover_comp(float *out, int comp, const Movers &M, size_t R, float q,
                                  const PP32 &P) {
  const MoverRec mr = M.rec[R];
  const int dy = mr.dir / 3 - 1, dx = mr.dir % 3 - 1;
  Move32 m;
  make_move32(m, mr.ox0, mr.oy0, mr.ox1, mr.oy1, dx, dy);
  const float qc = q * P.qc_scale;
  float acc[25];
#pragma unroll
  for (int v = 0; v < 25; ++v) acc[v] = 0.0f;
  if (comp == 0) {
    stencil32<0>(acc, m, qc, 0.0f);
#pragma unroll
    for (int v = 0; v < 20; ++v) out[v] = acc[v];
  } else if (comp == 1) {
    stencil32<1>(acc, m, qc, 0.0f);
#pragma unroll
    for (int v = 0; v < 20; ++v) out[20 + v] = acc[v];
  } else {
    const float ux = mr.ux, uy = mr.uy, uz = mr.uz;
    const float qv = q * (uz * __builtin_amdgcn_rsqf(1.0f + (ux * ux + uy * uy + uz * uz) * P.inv_c2));
    stencil32<2>(acc, m, qc, qv);
#pragma unroll
    for (int v = 0; v < 25; ++v) out[40 + v] = acc[v];
  }
}

// Push + deposit in place on the cell bins.  A particle that stays in its
// cell is written back to its slot; one that leaves goes to the cell's mover
// region (new state, old offsets for the deposit, ID) and leaves a hole.
// After the deposit of the movers, each bin closes its holes with its last
// particles (the k-th hole below the new count takes the k-th remaining
// particle above it), and records how many stayed and how many left in each
// of the 8 directions; wm_k_insert_movers then appends the movers to their
// destination bins.
__global__ __launch_bounds__(256, 4) void k_push_deposit32(
    PP32 P, int CB, const int *__restrict__ cnt, const float *__restrict__ tmp, int TX,
    size_t tps, Rec32 A, int64_t *__restrict__ ids, Movers M, int *__restrict__ ncnt,
    int *__restrict__ mcnt, double *__restrict__ part, int *__restrict__ derr) {
  __shared__ float T[6 * 3 * TW];
  __shared__ float4 CF4[CHUNK * CFS / 4];
  __shared__ double cellbuf[CHUNK * NACC];
  __shared__ int xl[CHUNK * WM_MG];                  // species << 28 | dir << 20 | slot
  __shared__ int pre[CHUNK + 1];                     // prefix of the mover counts
  __shared__ int gcnt[CHUNK];
  __shared__ float stg[XR * NACC];                   // stencils of one round of movers
  const int chunk = blockIdx.x;
  const int jl = blockIdx.y;
  const int i0 = chunk * CHUNK;                      // cell index relative to nxgs
  const int ncell = min(CHUNK, P.nx - i0);
  const size_t rbase = ((size_t)jl * P.nchunk + chunk) * CHUNK;   // first region
  for (int t = threadIdx.x; t < 6 * 3 * TW; t += blockDim.x) {
    const int col = t % TW, b = (t / TW) % 3, k = t / (3 * TW);
    T[t] = col < ncell + 2 ? tmp[(size_t)k * tps + (size_t)(i0 + col) + (size_t)TX * (jl + b)] : 0.0f;
  }
  for (int e = threadIdx.x; e < CHUNK * NACC; e += blockDim.x) cellbuf[e] = 0.0;
  __syncthreads();
  if (threadIdx.x < CHUNK * 6) {
    const int c = threadIdx.x / 6, k = threadIdx.x % 6;
    float *cf = (float *)CF4 + c * CFS;
    cell_coeffs(cf, T, c, k);
    if (k == 0) cf[54] = cf[55] = 0.0f;
  }
  __syncthreads();
  const int g = threadIdx.x / GROUP;
  const int sub = threadIdx.x % GROUP;
  const int gshift = GROUP * (g & 3);                // the group's lanes in the wave ballot
  const int il = i0 + g;                             // this group's cell
  double *cb = cellbuf + g * NACC;
  int *xg = xl + g * WM_MG;
  const size_t reg = (rbase + g) * WM_MG;            // this cell's mover region
  float core[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) core[k] = 0.0f;
  bool ok = true;
  int xcnt = 0;                                      // movers of the cell (all species)
  if (g < ncell) {
    for (int s = 0; s < P.nsp; ++s) {
      const size_t bin = bin_of(s, jl, il, P.ny, P.nx);
      const int n = min(CB, cnt[bin]);
      const size_t base = bin * (size_t)CB;
      // bin bases: 32-bit indices below
      float *Aox = A.ox + base, *Aoy = A.oy + base, *Aux = A.ux + base;
      float *Auy = A.uy + base, *Auz = A.uz + base;
      const float q = P.q[s], r = P.r[s];
      const float fac1 = q / r * 0.5f * P.delt;
      const float txxx = fac1 * fac1;
      const float fac2 = q * P.delt / r;
      const float qc = q * P.qc_scale;
      float nox = 0, noy = 0, nux = 0, nuy = 0, nuz = 0;
      if (sub < n) {
        nox = Aox[sub]; noy = Aoy[sub]; nux = Aux[sub]; nuy = Auy[sub]; nuz = Auz[sub];
      }
      // the window loop is uniform over the group, so every lane keeps the
      // same mover count
      for (int w = 0; w < n; w += GROUP) {
        const int k = w + sub;
        const float ox0 = nox, oy0 = noy, cux = nux, cuy = nuy, cuz = nuz;
        if (k + GROUP < n) {
          const int kn = k + GROUP;
          nox = Aox[kn]; noy = Aoy[kn]; nux = Aux[kn]; nuy = Auy[kn]; nuz = Auz[kn];
        }
        const bool on = k < n;
        bool cross = false;
        int dx = 0, dy = 0;
        float ox1 = 0.0f, oy1 = 0.0f;
        Push32 o{};
        if (on) {
          __asm__ volatile("" ::: "memory");
          o = boris32(CF4 + g * (CFS / 4), ox0, oy0, cux, cuy, cuz, P, fac1, txxx, fac2);
          dx = rebase(o.ox, ox1);
          dy = rebase(o.oy, oy1);
          ok = ok && fabsf(o.ox - 0.5f) < 1.5f && fabsf(o.oy - 0.5f) < 1.5f;
          cross = (dx | dy) != 0;
          if (!cross) {
            Aox[k] = ox1;
            Aoy[k] = oy1;
            Aux[k] = o.ux;
            Auy[k] = o.uy;
            Auz[k] = o.uz;
            const float qv =
                q * (o.uz * __builtin_amdgcn_rsqf(1.0f + (o.ux * o.ux + o.uy * o.uy + o.uz * o.uz) * P.inv_c2));
            core32(core, ox0, oy0, ox1, oy1, qc, qv);
          }
        }
        // movers in (species, slot) order
        const unsigned gm = (unsigned)(__ballot(cross) >> gshift) & 0xFFFFu;
        if (cross) {
          const int pos = xcnt + __popc(gm & ((1u << sub) - 1u));
          if (pos < WM_MG) {
            const size_t R = reg + pos;
            int ci = il + dx;
            ci = ci < 0 ? ci + P.nx : (ci >= P.nx ? ci - P.nx : ci);   // periodic x
            int cj = jl + dy;
            cj = cj < 0 ? cj + P.ny : (cj >= P.ny ? cj - P.ny : cj);   // periodic y
            MoverRec mr;
            mr.dst = (int)bin_of(s, cj, ci, P.ny, P.nx);
            mr.dir = (dy + 1) * 3 + (dx + 1);
            mr.ox0 = ox0;
            mr.oy0 = oy0;
            mr.ox1 = ox1;
            mr.oy1 = oy1;
            mr.ux = o.ux;
            mr.uy = o.uy;
            mr.uz = o.uz;
            mr.pad = 0;
            mr.id = ids[base + k];
            M.rec[R] = mr;
            xg[pos] = (s << 28) | (((dy + 1) * 3 + (dx + 1)) << 20) | k;
          }
        }
        xcnt += __popc(gm);
      }
    }
    // stay-in-cell stencils of every species, summed over the group once
#pragma unroll
    for (int c = 0; c < 27; ++c) {
      const float v = row_sum16(core[c]);
      if (sub == 0) cb[core_slot(c)] += (double)v;
    }
  }
  if (!ok) atomicOr(derr, DERR_JUMP);
  if (xcnt > WM_MG && sub == 0) atomicOr(derr, DERR_MOVER_OVER);
  if (sub == 0 && g < CHUNK) {
    gcnt[g] = g < ncell ? min(xcnt, WM_MG) : 0;
    M.n[rbase + g] = gcnt[g];
  }
  __syncthreads();                                   // also orders the movers' stores
  if (threadIdx.x == 0) {
    pre[0] = 0;
    for (int c = 0; c < CHUNK; ++c) pre[c + 1] = pre[c] + gcnt[c];
  }
  __syncthreads();
  // Movers of the whole workgroup: wave w computes stencil component w of
  // XR movers at a time (one mover per lane), then each (cell, value) pair
  // adds its cell's movers in list order.
  const int nxs = pre[CHUNK];
  for (int r0 = 0; r0 < nxs; r0 += XR) {
    const int comp = threadIdx.x / XR, e = r0 + threadIdx.x % XR;
    if (comp < 3 && e < nxs) {
      int gg = 0;
      while (pre[gg + 1] <= e) ++gg;
      const int j = e - pre[gg];
      const int s = xl[gg * WM_MG + j] >> 28;
      mover_comp(stg + (e - r0) * NACC, comp, M, (rbase + gg) * WM_MG + j, P.q[s], P);
    }
    __syncthreads();
    for (int pr = threadIdx.x; pr < ncell * NACC; pr += blockDim.x) {
      const int gg = pr / NACC, v = pr % NACC;
      const int lo = max(pre[gg], r0), hi = min(pre[gg + 1], r0 + XR);
      double a = cellbuf[pr];
      for (int e2 = lo; e2 < hi; ++e2) a += (double)stg[(e2 - r0) * NACC + v];
      cellbuf[pr] = a;
    }
    __syncthreads();
  }
  // close the holes and record the counts, one bin per (group, species)
  if (g < ncell) {
    const int m_all = gcnt[g];
    int j0 = 0;
    for (int s = 0; s < P.nsp; ++s) {
      int m = 0;
      while (j0 + m < m_all && (xg[j0 + m] >> 28) == s) ++m;
      const size_t bin = bin_of(s, jl, il, P.ny, P.nx);
      const int n = min(CB, cnt[bin]);
      const int n1 = n - m;
      const size_t base = bin * (size_t)CB;
      int a = 0;                                     // holes below n1
      while (a < m && (xg[j0 + a] & 0xFFFFF) < n1) ++a;
      for (int t = sub; t < a; t += GROUP) {
        // the t-th particle of [n1, n) that is not a hole
        int idx = a, c = 0, from = n1;
        for (; from < n; ++from) {
          if (idx < m && (xg[j0 + idx] & 0xFFFFF) == from) {
            ++idx;
            continue;
          }
          if (c == t) break;
          ++c;
        }
        if (from >= n) break;                          // (cannot happen: a tail particles)
        const size_t h = base + (xg[j0 + t] & 0xFFFFF), f = base + from;
        A.ox[h] = A.ox[f];
        A.oy[h] = A.oy[f];
        A.ux[h] = A.ux[f];
        A.uy[h] = A.uy[f];
        A.uz[h] = A.uz[f];
        ids[h] = ids[f];
      }
      if (sub < 8) {
        const int d = sub < 4 ? sub : sub + 1;       // directions 0..8 without 4
        int c = 0;
        for (int q = 0; q < m; ++q) c += ((xg[j0 + q] >> 20) & 0xF) == d;
        mcnt[bin * 8 + sub] = c;
      }
      if (sub == 0) ncnt[bin] = n1;
      j0 += m;
    }
  }
  double *out = part + ((size_t)jl * P.nchunk + chunk) * TILE;
  for (int e = threadIdx.x; e < TILE; e += blockDim.x) {
    const int comp = e % 3;
    const int t = (e / 3) % DW;
    const int b = e / (3 * DW);
    double sum = 0.0;
    const int glo = max(0, t - 4), ghi = min(ncell - 1, t);
    for (int gg = glo; gg <= ghi; ++gg) sum += acc_at(cellbuf + gg * NACC, comp, t - gg, b);
    out[e] = sum;
  }
}

// FP32 cell-centred fields (particle.f90:68-81 averages, in FP64, rounded).
__global__ void k_cell_fields32(FieldGeom fg, const double *__restrict__ uf,
                                float *__restrict__ tmp, int TX, int TY, size_t tps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= TX * TY) return;
  const int a = t % TX, b = t / TX;
  const size_t c0 = (size_t)(a + 1) + (size_t)fg.gx * (size_t)(b + 1);
  const size_t gx = fg.gx;
  const double *B1 = uf, *B2 = uf + fg.ps, *B3 = uf + 2 * fg.ps;
  const double *E1 = uf + 3 * fg.ps, *E2 = uf + 4 * fg.ps, *E3 = uf + 5 * fg.ps;
  tmp[t] = (float)(0.5 * (B1[c0] + B1[c0 + gx]));
  tmp[tps + t] = (float)(0.5 * (B2[c0] + B2[c0 + 1]));
  tmp[2 * tps + t] = (float)(0.25 * (B3[c0] + B3[c0 + 1] + B3[c0 + gx] + B3[c0 + gx + 1]));
  tmp[3 * tps + t] = (float)(0.5 * (E1[c0] + E1[c0 + 1]));
  tmp[4 * tps + t] = (float)(0.5 * (E2[c0] + E2[c0 + gx]));
  tmp[5 * tps + t] = (float)E3[c0];
}

// Sum of the FP32 kernel's tiles into uj (ghost-inclusive), fixed order.
__global__ void k_deposit_gather32(int nx, int ny, int nchunk, FieldGeom fg,
                                   const double *__restrict__ part, double *__restrict__ uj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= fg.gx * fg.gy) return;
  const int li = t % fg.gx, lj = t / fg.gx;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int chi = min(li / CHUNK, nchunk - 1);
  for (int b = 0; b < 5; ++b) {
    const int jl = lj - b;
    if (jl < 0 || jl >= ny) continue;
    for (int c = max(0, chi - 1); c <= chi; ++c) {
      const int tc = li - CHUNK * c;
      if (tc < 0 || tc >= DW) continue;
      const double *tp = part + ((size_t)jl * nchunk + c) * TILE + ((size_t)b * DW + tc) * 3;
      s0 += tp[0];
      s1 += tp[1];
      s2 += tp[2];
    }
  }
  (void)nx;
  uj[t] = s0;
  uj[fg.ps + t] = s1;
  uj[2 * fg.ps + t] = s2;
}

// ------------------------------------------ conversion to and from FP64
// Reference-layout staging rows (seg, k) of ndim doubles <-> Rec32 banks:
// ci = int(x) - nxgs, ox = x - int(x) rounded to FP32 (kept below 1), and
// likewise oy against the row j.
__global__ void k_aos_to_rec32(int ndim, int w, int cap, int ny, int nxgs, int nys,
                               const int *__restrict__ cnt, const double *__restrict__ st,
                               Rec32 B, int64_t *id) {
  const int seg = blockIdx.y;
  const int n = cnt[seg];
  const int j = nys + seg % ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double *r = st + ((size_t)seg * w + k) * ndim;
    const size_t p = (size_t)seg * cap + k;
    const int ix = (int)r[0];
    B.ci[p] = ix - nxgs;
    B.ox[p] = wm_below_one((float)(r[0] - ix));
    B.oy[p] = wm_below_one((float)(r[1] - j));
    B.ux[p] = (float)r[2];
    B.uy[p] = (float)r[3];
    B.uz[p] = (float)r[4];
    if (ndim >= 6) id[p] = __double_as_longlong(r[5]);
  }
}
__global__ void k_rec32_to_aos(int ndim, int w, int cap, int ny, int nxgs, int nys,
                               const int *__restrict__ cnt, double *__restrict__ st, Rec32 B,
                               const int64_t *id) {
  const int seg = blockIdx.y;
  const int n = cnt[seg];
  const int j = nys + seg % ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double *r = st + ((size_t)seg * w + k) * ndim;
    const size_t p = (size_t)seg * cap + k;
    r[0] = (double)(B.ci[p] + nxgs) + (double)B.ox[p];
    r[1] = (double)j + (double)B.oy[p];
    r[2] = B.ux[p];
    r[3] = B.uy[p];
    r[4] = B.uz[p];
    if (ndim >= 6) r[5] = __longlong_as_double(id[p]);
  }
}

// FP64 energy sums of FP32 momenta: per-workgroup partials.
// Workgroup (x, seg) sums the bins of cells x, x + nchunk, ... of segment seg.
__global__ __launch_bounds__(256) void k_energy_kin32(int nx, int CB, const int *cnt,
                                                      const float *ux, const float *uy,
                                                      const float *uz, double c, double *part,
                                                      int nchunk) {
  __shared__ double sh[256];
  const int seg = blockIdx.y;
  double v = 0.0;
  for (int i = blockIdx.x; i < nx; i += nchunk) {
    const size_t bin = (size_t)seg * nx + i;
    const int n = min(CB, cnt[bin]);
    for (int k = threadIdx.x; k < n; k += 256) {
      const size_t p = bin * CB + k;
      const double a = ux[p], b = uy[p], d = uz[p];
      v += sqrt(1 + (a * a + b * b + d * d) / (c * c)) - 1;
    }
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)seg * nchunk + blockIdx.x] = sh[0];
}

PP32 make_pp32(const wm_ctx *ctx) {
  PP32 P;
  const wm_grid &g = ctx->g;
  P.nx = g.nx; P.ny = g.ny; P.nxgs = g.nxgs; P.nys = g.nys; P.nsp = g.nsp;
  P.cap = ctx->cap;
  P.nchunk = wm_cdiv(g.nx, CHUNK);
  P.delt = (float)g.delt;
  P.c = (float)g.c;
  P.c2 = (float)(g.c * g.c);
  P.inv_c2 = (float)(1.0 / (g.c * g.c));
  P.qc_scale = (float)(g.delx / g.delt);
  for (int s = 0; s < WM_MAXSP; ++s) {
    P.q[s] = s < g.nsp ? (float)g.q[s] : 0.0f;
    P.r[s] = s < g.nsp ? (float)g.r[s] : 1.0f;
  }
  return P;
}

}  // namespace

size_t wm_deposit_part_elems32(int nx, int ny) {
  return (size_t)ny * (size_t)wm_cdiv(nx, CHUNK) * (size_t)TILE;
}

int wm_k_cell_fields32(wm_ctx *ctx) {
  const int TX = ctx->g.nx + 2, TY = ctx->g.ny + 2;
  const size_t tps = (size_t)TX * TY;
  hipLaunchKernelGGL(k_cell_fields32, dim3(wm_cdiv((long long)TX * TY, 256)), dim3(256), 0,
                     ctx->stream, ctx->fg, ctx->uf, ctx->tmp32, TX, TY, tps);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_push32_regions(int nx, int ny) { return ny * wm_cdiv(nx, CHUNK) * CHUNK; }

int wm_k_push_deposit32(wm_ctx *ctx) {
  PP32 P = make_pp32(ctx);
  const int TX = ctx->g.nx + 2;
  const size_t tps = (size_t)TX * (size_t)(ctx->g.ny + 2);
  hipLaunchKernelGGL(k_push_deposit32, dim3(P.nchunk, P.ny), dim3(256), 0, ctx->stream, P,
                     ctx->CB, ctx->bin_cnt, ctx->tmp32, TX, tps, ctx->b32[0], ctx->id_bins,
                     ctx->mv, ctx->bin_ncnt, ctx->bin_mcnt, ctx->dep_part, ctx->derr);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_deposit_gather32(wm_ctx *ctx) {
  const int n = ctx->fg.gx * ctx->fg.gy;
  hipLaunchKernelGGL(k_deposit_gather32, dim3(wm_cdiv(n, 256)), dim3(256), 0, ctx->stream,
                     ctx->g.nx, ctx->g.ny, wm_cdiv(ctx->g.nx, CHUNK), ctx->fg, ctx->dep_part,
                     ctx->uj);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_aos_rec32(wm_ctx *ctx, bool to_rec, int w, const int *dcnt, double *stage, int bank) {
  const wm_grid &g = ctx->g;
  const int nseg = g.nsp * g.ny;
  const dim3 grid(std::min(64, wm_cdiv(w, 256)), nseg);
  if (to_rec)
    hipLaunchKernelGGL(k_aos_to_rec32, grid, dim3(256), 0, ctx->stream, g.ndim, w, ctx->cap, g.ny,
                       g.nxgs, g.nys, dcnt, stage, ctx->b32[bank], ctx->id[ctx->idcur]);
  else
    hipLaunchKernelGGL(k_rec32_to_aos, grid, dim3(256), 0, ctx->stream, g.ndim, w, ctx->cap, g.ny,
                       g.nxgs, g.nys, dcnt, stage, ctx->b32[bank], ctx->id[ctx->idcur]);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_energy_kin32(wm_ctx *ctx, int bank, double *part, int nchunk) {
  const int nseg = ctx->g.nsp * ctx->g.ny;
  (void)bank;                                      // the bins are the state
  const Rec32 &B = ctx->b32[0];
  hipLaunchKernelGGL(k_energy_kin32, dim3(nchunk, nseg), dim3(256), 0, ctx->stream, ctx->g.nx,
                     ctx->CB, ctx->bin_cnt, B.ux, B.uy, B.uz, ctx->g.c, part, nchunk);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
