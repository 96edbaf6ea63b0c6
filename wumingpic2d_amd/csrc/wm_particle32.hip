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
#include "wm_migrate.h"

namespace {

constexpr int CHUNK = 16;             // cells per workgroup
constexpr int GROUP = 16;             // lanes per cell
constexpr int TW = CHUNK + 2;         // push tile width (halo 1)
constexpr int DW = CHUNK + 4;         // deposit tile width (halo 2)
constexpr int TILE = 5 * DW * 3;      // doubles per deposit tile
constexpr int NACC = 65;
constexpr int NT = 256;               // push workgroup size (a constant: blockDim.x is a
                                      // memory load, whose wait would include the stores)

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
  float ig;                   // 1 / gamma of the new momentum
};

// The TSC gather of particle.f90:107-129 inside one cell is a biquadratic
// in the offsets: with w-(o) = 1/2 - o + o^2/2, w0(o) = 1/2 + o - o^2,
// w+(o) = o^2/2, sum_ab T(a,b) w_a(ox) w_b(oy) = sum_ij C_ij ox^i oy^j.
// The workgroup computes C (in FP64, rounded once) for each of its cells and
// fields; a particle then needs 8 FMAs per field.
// Per cell: field k's 9 coefficients at floats 12 k .. 12 k + 8 (padded to
// 16 B), so a particle reads one field at a time.
constexpr int CFS = 72;
__device__ inline int coeff_at(int k) { return 12 * k; }
__device__ inline void cell_coeffs(float *cf, const float *T, int c, int k) {
  const double M[3][3] = {{0.5, -1.0, 0.5}, {0.5, 1.0, -1.0}, {0.0, 0.0, 0.5}};
  const float *Tk = T + k * 3 * TW + c;
  // C_ij = sum_b (sum_a T(b,a) M[a][i]) M[b][j], in FP64, rounded once
  double R[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      R[b][i] = (double)Tk[b * TW] * M[0][i] + (double)Tk[b * TW + 1] * M[1][i] +
                (double)Tk[b * TW + 2] * M[2][i];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      cf[coeff_at(k) + i * 3 + j] = (float)(R[0][i] * M[0][j] + R[1][i] * M[1][j] + R[2][i] * M[2][j]);
}

// particle.f90:96-161 in FP32 on the cell frame.
__device__ inline Push32 boris32(const float4 *cf4, float ox, float oy, float ux, float uy,
                                 float uz, const PP32 &P, float fac1, float txxx, float fac2) {
  float f[6];
  // one field at a time, E first: the memory clobber keeps the next field's
  // LDS reads after this one's use, so only 12 coefficients are live
#pragma unroll
  for (int kk = 0; kk < 6; ++kk) {
    const int k = (kk + 3) % 6;
    float C[12];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float4 v = cf4[3 * k + q];
      C[4 * q] = v.x; C[4 * q + 1] = v.y; C[4 * q + 2] = v.z; C[4 * q + 3] = v.w;
    }
    const float r0 = C[0] + oy * (C[1] + oy * C[2]);
    const float r1 = C[3] + oy * (C[4] + oy * C[5]);
    const float r2 = C[6] + oy * (C[7] + oy * C[8]);
    f[k] = r0 + ox * (r1 + ox * r2);
    __asm__ volatile("" ::: "memory");
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
  o.ig = ig;
  o.ox = ox + o.ux * P.delt * ig;
  o.oy = oy + o.uy * P.delt * ig;
  return o;
}

// Cell change and the offset rebased on the new cell, kept in [0, 1).
// Branch-free: o - d is exact for d = 1 (o in [1, 2)) and o itself for d = 0;
// for d = -1 the sum o + 1 may round up to 1, which the min takes below one.
__device__ inline int rebase(float o, float &onew) {
  const int d = (int)(o >= 1.0f) - (int)(o < 0.0f);
  onew = fminf(o - (float)d, 0x1.fffffep-1f);
  return d;
}

// Quadratic spline weights at the cell's -1, 0, +1 neighbours for the old
// offset o0, s- = (1 - o0)^2 / 2, s+ = o0^2 / 2, s0 = 1 - s- - s+, and their
// change to the new offset o1 of the same cell, taken directly rather than as
// a difference: d+ = (o1 - o0)(o1 + o0) / 2, d- = (o1 - o0)(o1 + o0 - 2) / 2,
// d0 = -(d- + d+).
__device__ inline void shape_delta(float o0, float o1, float *s, float *d) {
  const float t = 1.0f - o0;
  s[0] = (0.5f * t) * t;
  s[2] = (0.5f * o0) * o0;
  s[1] = (1.0f - s[0]) - s[2];
  const float hd = 0.5f * (o1 - o0), sg = o1 + o0;
  d[2] = hd * sg;
  d[0] = hd * (sg - 2.0f);
  d[1] = -(d[0] + d[2]);
}

// Stay-in-cell particle: the 21 possibly non-zero stencil values.  Of the
// 3x3 core, Jx's last row and Jy's last column vanish: their weights are
// -qc times the sum of all three shape changes, which is zero because the
// quadratic shape functions sum to one; for the same reason the middle row's
// weight -qc (d- + d0) is qc d+.  Jz's weight
// s0x s0y + dx s0y / 2 + s0x dy / 2 + dx dy / 3 is X Y + dx dy / 12 with
// X = s0x + dx / 2, Y = s0y + dy / 2.  Layout: Jx (a < 2, b < 3) at 0..5,
// Jy (a < 3, b < 2) at 6..11, Jz (3 x 3) at 12..20.
constexpr int NCORE = 21;
__device__ inline void core32(float *core, float ox0, float oy0, float ox1, float oy1, float qc,
                              float qv) {
  float sx[3], sy[3], dx[3], dy[3];
  shape_delta(ox0, ox1, sx, dx);
  shape_delta(oy0, oy1, sy, dy);
  float X[3], Y[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] = fmaf(0.5f, dx[k], sx[k]);
    Y[k] = fmaf(0.5f, dy[k], sy[k]);
  }
  const float wx0 = -(qc * dx[0]), wx1 = qc * dx[2];
  const float wy0 = -(qc * dy[0]), wy1 = qc * dy[2];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    core[b] = fmaf(wx0, Y[b], core[b]);
    core[3 + b] = fmaf(wx1, Y[b], core[3 + b]);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    core[6 + 2 * a] = fmaf(X[a], wy0, core[6 + 2 * a]);
    core[7 + 2 * a] = fmaf(X[a], wy1, core[7 + 2 * a]);
  }
  const float qv12 = qv * (1.0f / 12.0f);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float za = qv * X[a], zb = qv12 * dx[a];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float &c = core[12 + a * 3 + b];
      c = fmaf(za, Y[b], c);
      c = fmaf(zb, dy[b], c);
    }
  }
}

// A cell's stencil sums: the movers' sums, 65 per cell (Jx 0..19 as
// (a-1)*5 + b with tile offset a >= 1, Jy 20..39 as 20 + a*4 + (b-1) with
// b >= 1, Jz 40..64 as 40 + a*5 + b), and the stayers' 21 core sums where
// the core has the offset (Jx a = 2..3, Jy b = 2..3, offsets 1..3
// otherwise; the layout of core32).  Both are FP32 sums of the group's
// lanes; after the movers' deposit each group adds its cell's pair in FP64
// (csum, which takes the place of the movers' list offsets in LDS) for the
// tile assembly.
constexpr int CCS = 24;               // floats per cell of the core sums (21, padded)
constexpr int CMS = 68;               // floats per cell of the movers' sums (65, padded)
__shared__ float ccore[CHUNK * CCS];
__shared__ float cmov[CHUNK * CMS];
__device__ inline int core_of(int comp, int a, int b) {
  if (comp == 0) return (a >= 2 && a <= 3 && b >= 1 && b <= 3) ? (a - 2) * 3 + (b - 1) : -1;
  if (comp == 1) return (a >= 1 && a <= 3 && b >= 2 && b <= 3) ? 6 + (a - 1) * 2 + (b - 2) : -1;
  return (a >= 1 && a <= 3 && b >= 1 && b <= 3) ? 12 + (a - 1) * 3 + (b - 1) : -1;
}
// Entry e of the 65 sums in FP64: core + movers (the core term first).
__device__ inline double cell_sum(int cell, int e) {
  const int comp = e < 20 ? 0 : (e < 40 ? 1 : 2);
  const int a = comp == 0 ? e / 5 + 1 : (comp == 1 ? (e - 20) / 4 : (e - 40) / 5);
  const int b = comp == 0 ? e % 5 : (comp == 1 ? (e - 20) % 4 + 1 : (e - 40) % 5);
  const int c = core_of(comp, a, b);
  const double mv = (double)cmov[cell * CMS + e];
  const double cv = (double)ccore[cell * CCS + (c < 0 ? 0 : c)];
  return c < 0 ? mv : cv + mv;
}
// Sum of component comp at tile offset (a, b) of a cell, from csum.
__device__ inline double acc_at(const double *csum, int cell, int comp, int a, int b) {
  const int idx = comp == 0 ? (a - 1) * 5 + b : (comp == 1 ? 20 + a * 4 + (b - 1) : 40 + a * 5 + b);
  const bool none = (comp == 0 && a == 0) || (comp == 1 && b == 0);
  const double v = csum[cell * NACC + (none ? 0 : idx)];
  return none ? 0.0 : v;
}

// Sum over the 16 lanes of one DPP row (a group): quad butterflies, then
// rotations by 4 and 8 within the row.  Every lane ends with the total; the
// callers use lane 0's, whose association order is fixed.  The callers run
// it with every lane of the row active, so each DPP source lane is valid;
// written as update_dpp with bound_ctrl (zero for an invalid source), each
// step compiles to a single v_add_f32_dpp.
template <int CTRL>
__device__ inline float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ inline float row_sum16(float v) {
  v = dpp_add<0xB1>(v);         // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);         // quad_perm [2,3,0,1]
  v = dpp_add<0x124>(v);        // row_ror 4
  v = dpp_add<0x128>(v);        // row_ror 8
  __asm__ volatile("" : "+v"(v));   // the last add stays fused, not sunk into the caller's branch
  return v;
}




// Bin of (species s, row jl, cell i).
__device__ inline size_t bin_of(int s, int jl, int i, int ny, int nx) {
  return ((size_t)s * ny + jl) * nx + i;
}

// The movers of one push workgroup are listed, per cell in (species, slot)
// order, for their deposit, their bins' summaries and their records.  A
// cell's first CAPG entries sit in its own part of an LDS pool; a dense cell
// continues its list in blocks of MB entries of the row's spill pool in
// device memory, which its group takes as it reaches them (entry index
// NPOOL + b * MB + r is entry r of the row's spill block b).  Entry j of cell
// c's list is xent(c, j).  Which spill block a cell gets varies from run to
// run; what a list holds, and so every result, does not.
constexpr int MB = 16;                // entries per spill block
constexpr int CAPG = 60;              // entries of one cell's list in LDS
constexpr int NPOOL = CHUNK * CAPG;   // movers of one workgroup listed in LDS
constexpr int MAXB = 64;              // spill blocks of one cell's list
constexpr int HCN = NPOOL * 4 / (NT / 8);   // migration prologue: holes cached per bin group

// The push kernel's LDS (file scope, so device functions address it directly).
__shared__ float T[6 * 3 * TW];                   // cell-centred fields of the chunk
__shared__ float4 CF4[CHUNK * CFS / 4];           // per-cell gather coefficients
__shared__ int xl[NPOOL];                         // movers: species << 28 | dir << 20 | slot
__shared__ float4 xpl[NPOOL];                     // movers' offsets before and after
__shared__ float xql[NPOOL];                      // movers' charge times v_z
__shared__ uint16_t xblk[CHUNK * MAXB];           // spill block of each block of a cell's
                                                  // list past CAPG
__shared__ int gcnt[CHUNK];                       // movers listed per cell
__shared__ int scnt[WM_MAXSP * CHUNK];            // particles per (species, cell)
__shared__ int sbin[WM_MAXSP * CHUNK];            // first slot of each (species, cell) bin
__shared__ int gsb[WM_MAXSP * CHUNK];             // movers listed per cell up to the end of
                                                  // each species (the lists' species bounds)
__shared__ int rfirst;                            // first record of the workgroup's movers in
                                                  // its row's pool (-1: the pool is full)
__shared__ float sqc[WM_MAXSP];                   // q delx / delt per species (an LDS read:
                                                  // a per-lane index into the kernel arguments
                                                  // is a memory load, and its wait would also
                                                  // wait for the stores in flight)

// The row's spill of the mover lists: entries at base + (e - NPOOL).
struct Spill {
  int *l;
  float4 *p;
  float *q;
  size_t base;
  int *fill;                                      // the row's spill blocks used
  int nblk;                                       // spill blocks per row
};

__device__ inline int xent(int c, int j) {
  return j < CAPG ? c * CAPG + j
                  : NPOOL + (int)xblk[c * MAXB + ((j - CAPG) >> 4)] * MB + ((j - CAPG) & (MB - 1));
}
__device__ inline int xl_at(int e, const Spill &X) {
  return e < NPOOL ? xl[e] : X.l[X.base + (e - NPOOL)];
}
__device__ inline float4 xpl_at(int e, const Spill &X) {
  return e < NPOOL ? xpl[e] : X.p[X.base + (e - NPOOL)];
}
__device__ inline float xql_at(int e, const Spill &X) {
  return e < NPOOL ? xql[e] : X.q[X.base + (e - NPOOL)];
}

// The deposit of one cell's movers j = p, p + MQ, ... (part p of MQ) into
// the 65 stencil sums of the cell (Jx a < 4 at a*5 + b, Jy b < 4 at
// 20 + a*4 + b, Jz at 40 + a*5 + b): field.f90:224-268's Esirkepov weights
// with s0 the old and ds the new minus the old 5-point shapes, in FP32,
// summed over the part's movers in FP32.  Jz uses
// s0x s0y + ds_x s0y / 2 + s0x ds_y / 2 + ds_x ds_y / 3 = X Y + ds_x ds_y / 12
// with X = s0x + ds_x / 2, Y = s0y + ds_y / 2.
constexpr int MQ = 16;
// The 5-point old shape (support at 1..3) and new shape (support at
// 1 + inc .. 3 + inc): ds = new - old and X = old + ds / 2.  Selects on
// three flags rather than on inc's value, which the compiler would turn into
// a branching switch.
__device__ inline void place5(float sm, float s0, float sp, float nm, float n0, float np, int inc,
                              float *X, float *ds) {
  const bool lo = inc < 0, mid = inc == 0, hi = inc > 0;
  const float n[5] = {lo ? nm : 0.0f, lo ? n0 : (mid ? nm : 0.0f), lo ? np : (mid ? n0 : nm),
                      hi ? n0 : (mid ? np : 0.0f), hi ? np : 0.0f};
  const float o[5] = {0.0f, sm, s0, sp, 0.0f};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    ds[k] = n[k] - o[k];
    X[k] = o[k] + 0.5f * ds[k];
  }
}
__device__ inline void mover_part(float *acc, int p, int nm, int mc, const Spill &X) {
  const float twelfth = 1.0f / 12.0f;
  for (int j = p; j < nm; j += MQ) {
    const int e = xent(mc, j);
    const float4 v = xpl_at(e, X);
    const float qv = xql_at(e, X);
    const int xv = xl_at(e, X);
    const int dir = (xv >> 20) & 0xF;
    const int dy = dir / 3 - 1, dx = dir - 3 * (dy + 1) - 1;
    const float qc = sqc[xv >> 28];
    float xm, x0, xp_, ym, y0, yp, nxm, nx0, nxp, nym, ny0, nyp;
    shapef(v.x - 0.5f, xm, x0, xp_);
    shapef(v.y - 0.5f, ym, y0, yp);
    shapef(v.z - 0.5f, nxm, nx0, nxp);
    shapef(v.w - 0.5f, nym, ny0, nyp);
    float X[5], dX[5], Y[5], dY[5];
    place5(xm, x0, xp_, nxm, nx0, nxp, dx, X, dX);
    place5(ym, y0, yp, nym, ny0, nyp, dy, Y, dY);
    // Jx (a, b) = w_a Y_b, w_a = -qc sum_{a' <= a} ds_x[a']; Jy likewise in y
    float wx = 0.0f, wy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) wy[k] = (k ? wy[k - 1] : 0.0f) - qc * dY[k];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      wx -= qc * dX[a];
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[a * 5 + b] += wx * Y[b];
    }
#pragma unroll
    for (int a = 0; a < 5; ++a) {
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[20 + a * 4 + b] += X[a] * wy[b];
      const float za = qv * X[a], zb = (qv * twelfth) * dX[a];
#pragma unroll
      for (int b = 0; b < 5; ++b) acc[40 + a * 5 + b] += za * Y[b] + zb * dY[b];
    }
  }
}

constexpr int PD = 5;           // windows of a bin held in registers (80 particles)

// The first PD windows of one bin, loaded in one burst.
struct Win {
  float ox[PD], oy[PD], ux[PD], uy[PD], uz[PD];
};

struct Spc {                    // one species of one cell
  int s, n, off;                // off: first slot of the group's bin in its segment
  size_t base0;                 // first slot of the species' row segment (workgroup-uniform)
  float q, fac1, txxx, fac2, qc;
};
constexpr int BUF_WORD3 = 0x00020000;   // raw (untyped) buffer access on gfx9-family GPUs
// Buffer resource of nbytes from p, which is the same in every lane (a row
// segment's start): read from the first lane into scalar registers, so that
// the buffer instructions take it directly.
__device__ inline __amdgpu_buffer_rsrc_t segment_rsrc(const void *p, int nbytes) {
  const unsigned long long a = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void *u = (void *)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(u, 0, nbytes, BUF_WORD3);
}
// First slot of the group's bin: the workgroup-uniform segment start plus the
// bin's 32-bit offset, so that no 64-bit per-lane base stays live through the
// windows.
__device__ inline size_t spc_base(const Spc &S) { return S.base0 + (size_t)(unsigned)S.off; }

__device__ inline void make_spc(Spc &S, int s, int n, int off, int jl, const PP32 &P,
                                const Bins32 &A) {
  S.s = s;
  S.n = n;
  S.off = off;
  S.base0 = ((size_t)s * P.ny + jl) * (size_t)P.cap;

  S.q = P.q[s];
  S.fac1 = P.q[s] / P.r[s] * 0.5f * P.delt;
  S.txxx = S.fac1 * S.fac1;
  S.fac2 = P.q[s] * P.delt / P.r[s];
  S.qc = P.q[s] * P.qc_scale;
}

// Window w of a bin into the register slot w.  The load is unconditional
// (a lane past the bin's count re-reads its last particle, or slot 0 of an
// empty bin): a conditional load would first clear the registers, and that
// write makes the compiler wait for every load still in flight.
__device__ inline void load_slot(Win &W, int w, const Bins32 &A, const Spc &S, int sub) {
  const size_t p = spc_base(S) + min(w * GROUP + sub, max(S.n - 1, 0));
  const float4 v = A.q[p];
  W.ox[w] = v.x;
  W.oy[w] = v.y;
  W.ux[w] = v.z;
  W.uy[w] = v.w;
  W.uz[w] = A.uz[p];
}

// Per-lane state of one group's pass over its cell; the group's indices
// follow from threadIdx and blockIdx.
struct Lane {
  int xcnt;                     // movers of the cell listed so far (all species)
  bool ok;                      // no particle jumped more than one cell
  bool lost;                    // a mover found no room in the pool
};
__device__ inline int lane_g() { return threadIdx.x / GROUP; }
__device__ inline int lane_sub() { return threadIdx.x % GROUP; }

// One window of GROUP particles (slot k = w * GROUP + sub of the bin); the
// group executes it together.  Every particle's new state is written back to
// its slot (a mover's slot becomes a hole, so what it holds does not matter;
// the window's stores stay contiguous).  A stayer's core stencil is
// accumulated; a mover is listed in LDS with the inputs of its deposit (its
// record is written after the loop, from its slot).
// The push of one window's particles: the new state, the stayers' stencils,
// the movers' list entries.
__device__ inline void push_compute(Lane &L, float *core, int k, bool on, float ox0, float oy0,
                                    float cux, float cuy, float cuz, const Spc &S, const PP32 &P,
                                    const Spill &X, float &ox1, float &oy1, float &nux, float &nuy,
                                    float &nuz) {
  __asm__ volatile("" ::: "memory");               // gather coefficients re-read from LDS
  const Push32 o = boris32(CF4 + lane_g() * (CFS / 4), ox0, oy0, cux, cuy, cuz, P, S.fac1, S.txxx,
                           S.fac2);
  const int dx = rebase(o.ox, ox1);
  const int dy = rebase(o.oy, oy1);
  nux = o.ux;
  nuy = o.uy;
  nuz = o.uz;
  L.ok = L.ok & (!on | (fmaxf(fabsf(o.ox - 0.5f), fabsf(o.oy - 0.5f)) < 1.5f));
  const bool cross = on && (dx | dy) != 0;
  const float qv = S.q * (o.uz * o.ig);            // charge times v_z
  if (on && !cross) {
    core32(core, ox0, oy0, ox1, oy1, S.qc, qv);
  }
  // movers in (species, slot) order: the window's movers are entries
  // pos0 .. pos0 + nmv - 1 of the cell's list
  const unsigned gm = (unsigned)(__ballot(cross) >> (GROUP * (lane_g() & 3))) & 0xFFFFu;
  const int nmv = __popc(gm);
  if (nmv != 0 && !L.lost) {                       // uniform over the group
    const int pos0 = L.xcnt;
    const int j = pos0 + __popc(gm & ((1u << lane_sub()) - 1u));
    bool room = true;
    if (pos0 + nmv > CAPG) {
      // the list runs past the cell's LDS entries: the group's lane 0 takes
      // the spill blocks that this window's movers start
      const int taken = pos0 > CAPG ? (pos0 - CAPG + MB - 1) / MB : 0;
      const int need = (pos0 + nmv - CAPG + MB - 1) / MB;
      const int nnew = need - taken;
      int first = 0;
      if (lane_sub() == 0 && nnew > 0) {
        const int f = atomicAdd(X.fill, nnew);
        first = f + nnew <= X.nblk ? f : -1;
      }
      first = __shfl(first, (int)(threadIdx.x & 63) & ~(GROUP - 1), 64);
      room = first >= 0 && need <= MAXB;
      if (room && lane_sub() < nnew) xblk[lane_g() * MAXB + taken + lane_sub()] = (uint16_t)(first + lane_sub());
      __builtin_amdgcn_wave_barrier();
    }
    if (cross && room) {
      const int e = xent(lane_g(), j);
      const int dir = (dy + 1) * 3 + (dx + 1);
      const int v = (S.s << 28) | (dir << 20) | k;
      const float4 pp = make_float4(ox0, oy0, ox1, oy1);
      if (e < NPOOL) {
        xl[e] = v;
        xpl[e] = pp;
        xql[e] = qv;
      } else {
        const size_t g = X.base + (e - NPOOL);
        X.l[g] = v;
        X.p[g] = pp;
        X.q[g] = qv;
      }
    }
    if (room)
      L.xcnt += nmv;
    else
      L.lost = true;
  }
}

__device__ inline void push_window(Lane &L, float *core, int k, float ox0, float oy0, float cux,
                                   float cuy, float cuz, const Spc &S, const PP32 &P,
                                   const Bins32 &A, const Spill &X) {
  // A window none of whose lanes holds a particle (the register windows past
  // the wave's fullest bin) skips the push; every window still issues its
  // stores, which the hardware drops for a lane past the bin's count (its
  // offset lies past the segment).  Stores outside any branch keep the
  // compiler's count of memory operations in flight exact, so its waits for
  // the register windows do not also wait for recent stores.
  const bool on = k < S.n;
  // (what a skipped window stores is dropped: no lane is on)
  float ox1 = 0.0f, oy1 = 0.0f, nux = 0.0f, nuy = 0.0f, nuz = 0.0f;
  if (__builtin_amdgcn_ballot_w64(on) != 0ULL)
    push_compute(L, core, k, on, ox0, oy0, cux, cuy, cuz, S, P, X, ox1, oy1, nux, nuy, nuz);
  // buffer resources of the segment's q and uz slots: a store at an offset
  // past the segment is dropped by the hardware
  const __amdgpu_buffer_rsrc_t rq = segment_rsrc(A.q + S.base0, P.cap * 16);
  const __amdgpu_buffer_rsrc_t ru = segment_rsrc(A.uz + S.base0, P.cap * 4);
  const unsigned sl = (unsigned)(S.off + k);
  const unsigned bq = on ? sl * 16u : 0xFFFFFFF0u, bu = on ? sl * 4u : 0xFFFFFFFCu;
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(nuz), ru, bu, 0, 0);
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 qv4 = {__float_as_uint(ox1), __float_as_uint(oy1), __float_as_uint(nux), __float_as_uint(nuy)};
  __builtin_amdgcn_raw_buffer_store_b128(qv4, rq, bq, 0, 0);
  __builtin_amdgcn_sched_barrier(0);               // windows are not interleaved (registers)
}


// All windows of one species of the group's cell: the PD register slots
// (all of them, whatever the count), each refilled with the same window of
// the next species (when there is one) as soon as it is pushed; then the rest
// of bins above PD * GROUP particles, loaded directly.
template <bool MORE>
__device__ inline void push_species(Lane &L, float *core, Win &W, const Spc &S, const Spc &Sn,
                                    const PP32 &P, const Bins32 &A, const Spill &X) {
#pragma unroll
  for (int w = 0; w < PD; ++w) {
    push_window(L, core, w * GROUP + lane_sub(), W.ox[w], W.oy[w], W.ux[w], W.uy[w], W.uz[w], S, P, A,
                X);
    if (MORE) load_slot(W, w, A, Sn, lane_sub());
  }
  // Bins above PD windows (dense cells of the nonlinear phase): each window's
  // load is issued before the window ahead of it is pushed.  A window's
  // stores touch only its own slots; the next window's loads (a lane past
  // the count re-reads the last slot, which lies in that next window) read
  // slots they do not.
  if (PD * GROUP < S.n) {
    int w = PD * GROUP;
    size_t p = spc_base(S) + min(w + lane_sub(), S.n - 1);
    float4 v = A.q[p];
    float u = A.uz[p];
    for (; w < S.n; w += GROUP) {
      const float4 cv = v;
      const float cu = u;
      p = spc_base(S) + min(w + GROUP + lane_sub(), S.n - 1);
      v = A.q[p];
      u = A.uz[p];
      push_window(L, core, w + lane_sub(), cv.x, cv.y, cv.z, cv.w, cu, S, P, A, X);
    }
    // nothing of this path stays in flight: otherwise the compiler, not
    // knowing whether the loop ran, would drain the register windows after it
    __builtin_amdgcn_s_waitcnt(0);
  }
}

// Push + deposit in place on the cell bins.  Every particle's new state is
// written to its slot; one that leaves its cell is listed in LDS and its slot
// becomes a hole.  After the loop the movers' deposit is summed per cell from
// the list, each mover's record (new state, source slot and direction, ID)
// goes to the row's mover pool, in a range the workgroup takes with one
// atomic, and each bin's summary records how many of its particles left, in
// total (with the index of its first record in the pool) and in each of the 8
// directions, and the slots and directions of the first WM_SUM_MV of them.
// The next push's prologue (or the settle kernel, or a re-layout of the row)
// then moves the movers into their destination bins.
// PAIR: exactly two species (straight-line code, so the compiler keeps the
// second species' windows in flight through the first without spilling).
template <bool PAIR>
__global__ __launch_bounds__(256, 4) void k_push_deposit32(
    PP32 P, const float *__restrict__ tmp, int TX, size_t tps, Bins32 A, int64_t *ids,
    wm_mig::Mig G, Movers Mp, const int *__restrict__ rowpend, Movers M, int *__restrict__ mfill,
    int mrow, Spill X, BinSum *__restrict__ msum, double *__restrict__ part,
    int *__restrict__ stats, int *__restrict__ derr) {
  const int chunk = blockIdx.x;
  const int jl = blockIdx.y;
  const int i0 = chunk * CHUNK;                      // cell index relative to nxgs
  const int ncell = min(CHUNK, P.nx - i0);
  X.base = (size_t)jl * (size_t)(X.nblk * MB);
  X.fill += jl;
  // the chunk's cell-centred fields first: they do not depend on the
  // migration, so their loads overlap its
  constexpr int NTV = (6 * 3 * TW + NT - 1) / NT;
  float tv[NTV];
#pragma unroll
  for (int q = 0; q < NTV; ++q) {
    const int t = threadIdx.x + q * NT;
    const int col = t % TW, b = (t / TW) % 3, k = t / (3 * TW);
    tv[q] = t < 6 * 3 * TW && col < ncell + 2
                ? tmp[(size_t)k * tps + (size_t)(i0 + col) + (size_t)TX * (jl + b)]
                : 0.0f;
  }
  // The last push's movers into this workgroup's bins (8 lanes per bin),
  // unless the row was laid out anew since (its bins hold them); they read
  // only what the last push wrote, and write only these bins, so workgroups
  // need not wait for one another.  Each bin's count and first slot go
  // straight to LDS; the barrier below also makes the bins' new slots visible
  // to the window loads.
  const int pend = rowpend[jl];
  for (int u = threadIdx.x; u < P.nsp * CHUNK * wm_mig::DL; u += NT) {
    const int k = u % wm_mig::DL, c = (u / wm_mig::DL) % CHUNK, s = u / (wm_mig::DL * CHUNK);
    int cn = 0, b0 = 0;
    if (c < ncell) {
      const size_t seg = (size_t)s * P.ny + jl;
      size_t base;
      int capd;
      wm_mig::bin_span(G, seg, i0 + c, base, capd);
      b0 = (int)(base - seg * (size_t)P.cap);
      const size_t d = seg * P.nx + i0 + c;
      // (the movers' offsets pool, unused until the windows, caches holes)
      int *hc = (int *)xpl + (threadIdx.x / wm_mig::DL) * HCN;
      cn = pend ? wm_mig::migrate_bin(G, Mp, A, ids, derr, d, i0 + c, jl, s, k, base, capd, hc, HCN)
                : min(capd, G.cnt[d]);
    }
    if (k == 0) {
      scnt[s * CHUNK + c] = cn;
      sbin[s * CHUNK + c] = b0;
    }
  }
  if (threadIdx.x < WM_MAXSP) sqc[threadIdx.x] = P.q[threadIdx.x] * P.qc_scale;
#pragma unroll
  for (int q = 0; q < NTV; ++q)
    if (threadIdx.x + q * NT < 6 * 3 * TW) T[threadIdx.x + q * NT] = tv[q];
  __syncthreads();
  const int g = threadIdx.x / GROUP;
  const int sub = threadIdx.x % GROUP;
  const int il = i0 + g;                             // this group's cell
  // the first species' windows are in flight during the coefficient set-up
  Win win;
  Spc sc{}, sn{};
  if (g < ncell) {
    make_spc(sc, 0, scnt[g], sbin[g], jl, P, A);
#pragma unroll
    for (int w = 0; w < PD; ++w) load_slot(win, w, A, sc, sub);
  }
  if (threadIdx.x < CHUNK * 6) {
    const int c = threadIdx.x / 6, k = threadIdx.x % 6;
    float *cf = (float *)CF4 + c * CFS;
    cell_coeffs(cf, T, c, k);
    cf[12 * k + 9] = cf[12 * k + 10] = cf[12 * k + 11] = 0.0f;
  }
  __syncthreads();
  Lane L;
  L.xcnt = 0;
  L.ok = true;
  L.lost = false;
  float core[NCORE];
#pragma unroll
  for (int k = 0; k < NCORE; ++k) core[k] = 0.0f;
  if (g < ncell) {
    // the next species' windows load while this one is pushed
    if (PAIR) {
      make_spc(sn, 1, scnt[CHUNK + g], sbin[CHUNK + g], jl, P, A);
      push_species<true>(L, core, win, sc, sn, P, A, X);
      const int x0 = L.xcnt;
      push_species<false>(L, core, win, sn, sn, P, A, X);
      if (sub == 0) {
        gsb[g] = x0;
        gsb[CHUNK + g] = L.xcnt;
      }
    } else {
      for (int s = 0; s + 1 < P.nsp; ++s) {
        make_spc(sn, s + 1, scnt[(s + 1) * CHUNK + g], sbin[(s + 1) * CHUNK + g], jl, P, A);
        push_species<true>(L, core, win, sc, sn, P, A, X);
        if (sub == 0) gsb[s * CHUNK + g] = L.xcnt;
        sc = sn;
      }
      push_species<false>(L, core, win, sc, sn, P, A, X);
      if (sub == 0) gsb[(P.nsp - 1) * CHUNK + g] = L.xcnt;
    }
    // stay-in-cell stencils of every species, summed over the group once
#pragma unroll
    for (int c = 0; c < NCORE; ++c) {
      const float v = row_sum16(core[c]);
      if (sub == 0) ccore[g * CCS + c] = v;
    }
  }
  if (!L.ok) atomicOr(derr, DERR_JUMP);
  if (L.lost && sub == 0) atomicOr(derr, DERR_MOVER_OVER);
  if (sub == 0 && g < CHUNK) gcnt[g] = g < ncell ? L.xcnt : 0;
  __syncthreads();                                   // the lists, and the movers' new state in
                                                     // their slots
  // the workgroup's records in the row's pool (one atomic), while every
  // cell's group deposits the cell's movers: lane sub takes movers j = sub,
  // sub + MQ, ... (a dense cell of the nonlinear phase loses a hundred
  // particles a step), and the group sums its lanes in a fixed order
  static_assert(CHUNK * MQ == NT && MQ == GROUP, "the movers' deposit runs on every cell group");
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int c = 0; c < CHUNK; ++c) tot += gcnt[c];
    int f = tot > 0 ? atomicAdd(mfill + jl, tot) : 0;
    if (f + tot > mrow) {
      atomicOr(derr, DERR_MOVER_OVER);
      f = -1;
    }
    rfirst = f;
    atomicMax(stats + 2, tot);
  }
  {
    float acc[NACC];
#pragma unroll
    for (int e = 0; e < NACC; ++e) acc[e] = 0.0f;
    mover_part(acc, sub, g < ncell ? gcnt[g] : 0, g, X);
    const bool lead = sub == 0 && g < ncell;
#pragma unroll
    for (int e = 0; e < NACC; ++e) {
      const float v = row_sum16(acc[e]);
      if (lead) cmov[g * CMS + e] = v;
    }
  }
  __syncthreads();
  const int rf = rfirst;
  double *csum = (double *)xpl;                      // the list offsets are read by now
  if (g < ncell) {
    int before = 0;
    for (int c = 0; c < g; ++c) before += gcnt[c];
    const int gm = gcnt[g];
    // each of the cell's movers' records, by its group: new state from its
    // slot, ID, source slot and direction
    for (int j = sub; j < gm && rf >= 0; j += GROUP) {
      const int xv = xl_at(xent(g, j), X);
      const int s = xv >> 28, dir = (xv >> 20) & 0xF, k = xv & 0xFFFFF;
      const size_t p = ((size_t)s * P.ny + jl) * (size_t)P.cap + (unsigned)sbin[s * CHUNK + g] + k;
      const float4 v = A.q[p];
      const float uz = A.uz[p];
      const int64_t id = ids ? ids[p] : 0;
      float4 *rp = (float4 *)(M.rec + (size_t)jl * mrow + rf + before + j);
      rp[0] = v;
      rp[1] = make_float4(uz, __int_as_float(k | ((dir - (dir > 4)) << 24)),
                          __int_as_float((int)(uint32_t)id), __int_as_float((int)(id >> 32)));
    }
    // the cell's sums in FP64 for the tile assembly
    for (int e = sub; e < NACC; e += GROUP) csum[g * NACC + e] = cell_sum(g, e);
    // per bin: movers in total and per direction
    for (int s = 0; s < P.nsp; ++s) {
      const int j0 = s ? gsb[(s - 1) * CHUNK + g] : 0;
      const int m = gsb[s * CHUNK + g] - j0;
      BinSum *bs = msum + bin_of(s, jl, il, P.ny, P.nx);
      // lane sub < 8 of the cell's group: direction index sub (directions
      // 0..8 without 4); its count, and (a bin with more movers than its
      // summary lists) the bin's records of that direction in list order at
      // perm[first + (counts of the directions before) + rank]
      const int d = sub < 4 ? sub : sub + 1;
      int c = 0;
      if (sub < 8)
        for (int q = 0; q < m; ++q) c += ((xl_at(xent(g, j0 + q), X) >> 20) & 0xF) == d;
      int pre = c;
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) {
        const int v = __shfl_up(pre, o, 8);
        if (sub >= o) pre += v;
      }
      pre -= c;
      if (sub < 8) {
        if (c > 255) atomicOr(derr, DERR_MOVER_OVER);
        bs->cnt[sub] = (uint8_t)min(c, 255);
        if (rf >= 0 && c > 0 && m > WM_SUM_MV) {       // else the summary lists them
          int *pp = M.perm + (size_t)jl * mrow + rf + before + j0 + pre;
          for (int q = 0, r = 0; r < c; ++q)
            if (((xl_at(xent(g, j0 + q), X) >> 20) & 0xF) == d) pp[r++] = q;
        }
      }
      if (sub < min(m, WM_SUM_MV)) {
        const int xv = xl_at(xent(g, j0 + sub), X);
        const int dir = (xv >> 20) & 0xF;
        bs->mv[sub] = (uint16_t)((xv & 0xFFFFF) | ((dir - (dir > 4)) << WM_SLOT_BITS));
      }
      if (sub == 0) {
        bs->first = (uint32_t)(max(rf, 0) + before + j0);
        bs->m = (uint32_t)m;
      }
    }
  }
  __syncthreads();                                   // every cell's FP64 sums
  double *out = part + ((size_t)jl * P.nchunk + chunk) * TILE;
  for (int e = threadIdx.x; e < TILE; e += NT) {
    const int comp = e % 3;
    const int t = (e / 3) % DW;
    const int b = e / (3 * DW);
    double sum = 0.0;
    const int glo = max(0, t - 4), ghi = min(ncell - 1, t);
    for (int gg = glo; gg <= ghi; ++gg) sum += acc_at(csum, gg, comp, t - gg, b);
    out[e] = sum;
  }
}

// FP32 cell-centred fields (particle.f90:68-81 averages, in FP64, rounded).
// Also clears the row pools' record counts and spill counts of the push that
// follows.
__global__ void k_cell_fields32(FieldGeom fg, const double *__restrict__ uf,
                                float *__restrict__ tmp, int TX, int TY, size_t tps,
                                int *__restrict__ mfill, int *__restrict__ sfill, int ny) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ny) {
    mfill[t] = 0;
    sfill[t] = 0;
  }
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
// Also marks every row's bins as awaiting the push's movers.
__global__ void k_deposit_gather32(int nx, int ny, int nchunk, FieldGeom fg,
                                   const double *__restrict__ part, double *__restrict__ uj,
                                   int *__restrict__ rowpend) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ny) rowpend[t] = 1;
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

// FP64 energy sums of FP32 momenta: per-workgroup partials.
// Workgroup (x, seg) sums the bins of cells x, x + nchunk, ... of segment seg,
// a wavefront per bin (wavefront w takes every fourth of them) and a lane
// per slot, in a fixed order.  Between a push and the next migration a
// mover's slot still holds its pushed state (its record is a copy), so the
// bins' slots below their counts hold every particle once.
__global__ __launch_bounds__(256) void k_energy_kin32(int nx, int cap, const int *bstart,
                                                      const int *cnt, const float4 *q,
                                                      const float *uz, double c, double *part,
                                                      int nchunk) {
  __shared__ double sh[4];
  const int seg = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v = 0.0;
  for (int i = blockIdx.x + nchunk * w; i < nx; i += 4 * nchunk) {
    const size_t bin = (size_t)seg * nx + i;
    const int n = cnt[bin];
    const size_t base = (size_t)seg * cap + bstart[(size_t)seg * (nx + 1) + i];
    for (int k = lane; k < n; k += 64) {
      const size_t p = base + k;
      const float4 t = q[p];
      const double a = t.z, b = t.w, d = uz[p];
      v += sqrt(1 + (a * a + b * b + d * d) / (c * c)) - 1;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)seg * nchunk + blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
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
  // the push that follows writes its movers to the other buffer when the
  // last push's are pending (wm_k_push_deposit32)
  const int out = ctx->mv_pending ? 1 - ctx->mv_last : ctx->mv_last;
  hipLaunchKernelGGL(k_cell_fields32, dim3(wm_cdiv((long long)std::max(TX * TY, ctx->g.ny), 256)),
                     dim3(256), 0, ctx->stream, ctx->fg, ctx->uf, ctx->tmp32, TX, TY, tps,
                     ctx->mfill[out], ctx->sfill, ctx->g.ny);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_push_deposit32(wm_ctx *ctx) {
  PP32 P = make_pp32(ctx);
  const int TX = ctx->g.nx + 2;
  const size_t tps = (size_t)TX * (size_t)(ctx->g.ny + 2);
  // the last push's movers (if not yet placed) are read from its buffer,
  // this push's go to the other
  const int in = ctx->mv_last, out = ctx->mv_pending ? 1 - in : in;
  const wm_mig::Mig G = wm_make_mig(ctx, in);
  hipLaunchKernelGGL(P.nsp == 2 ? k_push_deposit32<true> : k_push_deposit32<false>,
                     dim3(P.nchunk, P.ny), dim3(NT), 0, ctx->stream, P, ctx->tmp32, TX, tps,
                     ctx->bins, ctx->id_bins, G, ctx->mv[in], ctx->rowpend, ctx->mv[out],
                     ctx->mfill[out], ctx->mrow,
                     Spill{ctx->spill_l, ctx->spill_p, ctx->spill_q, 0, ctx->sfill, ctx->srow / MB},
                     ctx->bin_sum[out], ctx->dep_part, ctx->fstat, ctx->derr);
  WM_HIP(hipGetLastError());
  ctx->mv_last = out;
  ctx->mv_pending = true;
  return WM_OK;
}

int wm_push32_pool() { return NPOOL; }

int wm_k_deposit_gather32(wm_ctx *ctx) {
  const int n = std::max(ctx->fg.gx * ctx->fg.gy, ctx->g.ny);
  hipLaunchKernelGGL(k_deposit_gather32, dim3(wm_cdiv(n, 256)), dim3(256), 0, ctx->stream,
                     ctx->g.nx, ctx->g.ny, wm_cdiv(ctx->g.nx, CHUNK), ctx->fg, ctx->dep_part,
                     ctx->uj, ctx->rowpend);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_k_energy_kin32(wm_ctx *ctx, int bank, double *part, int nchunk) {
  const int nseg = ctx->g.nsp * ctx->g.ny;
  (void)bank;                                      // the bins are the state
  hipLaunchKernelGGL(k_energy_kin32, dim3(nchunk, nseg), dim3(256), 0, ctx->stream, ctx->g.nx,
                     ctx->cap, ctx->bstart, ctx->bin_cnt, ctx->bins.q, ctx->bins.uz, ctx->g.c, part,
                     nchunk);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
