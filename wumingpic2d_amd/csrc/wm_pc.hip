// wm_pc.hip -- learned preconditioner for the implicit field solve.
//
// The reference solves (f4 I - N4) dB_l = f5 gkl_l with plain CG
// (field.f90:319-461).  This file holds a four-level linear convolutional
// net M(r) ~ (f4 I - N4)^-1 r, applied to the three components as a batch of
// three images, and its training on (residual, correction) pairs that the
// plain solver records.  The solver side (flexible CG, fallback) lives with
// the CG kernels in wm_field.hip.
//
// Net (every convolution 3x3, periodic; level l has (nx >> l) x (ny >> l)
// pixels, channels C0 = C1 = 16, C2 = 32):
//   x0 = r / s                  s = rms(r) of the image
//   a0 = lift(x0)               1 -> C0, FP32, never stored
//   d1 = pool(a0)               2x2 mean                         BF16
//   a1 = D1(d1), d2 = pool(a1), a2 = D2(d2), d3 = pool(a2), a3 = D3(d3)
//   u2 = a2 + up(a3), b2 = U2(u2), u1 = a1 + up(b2), b1 = U1(u1)
//   u0 = a0 + up(b1)            FP32
//   z  = s * proj(u0)           C0 -> 1, FP32
// D1..U1 run on MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation) with BF16 activations in NHWC layout.  The fine level never
// touches HBM except for r and z: lift+pool is one kernel, and up+add+proj
// recomputes a0 from r in LDS.
//
// Training: L = mean((y - e/s)^2), gradients by hand-written backward
// kernels (adjoint convolutions reuse the MFMA kernel with flipped weights,
// weight gradients are per-workgroup partial sums added in a fixed order),
// Adam on FP32 master weights.  Everything is deterministic.
#include <math.h>

#include "wm_internal.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int C0 = 16, C1 = 16, C2 = 32;
constexpr int T = 16;     // tile edge (pixels) of every kernel

// Layers in the flat parameter vector, each [tap][cout][cin],
// tap = (dy+1)*3 + (dx+1).
enum { L_LIFT, L_D1, L_D2, L_D3, L_U2, L_U1, L_PROJ, NLAYER };
constexpr int LCIN[NLAYER] = {1, C0, C1, C2, C2, C1, C0};
constexpr int LCOUT[NLAYER] = {C0, C1, C2, C2, C1, C0, 1};
__host__ __device__ constexpr int loff(int l) {
  return l == 0 ? 0 : loff(l - 1) + 9 * LCIN[l - 1] * LCOUT[l - 1];
}
constexpr int NPARAM = loff(NLAYER);
constexpr int MAXW = 9 * 32 * 32;   // largest layer
constexpr int FINE_BLOCKS = 1024;   // workgroups of the fine-level weight gradients

__device__ inline int pmod(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

// Tile of this workgroup in a (tiles x, tiles y, images) grid, renumbered so
// that each XCD sweeps a contiguous band of tiles.  The dispatcher deals
// workgroups round-robin to the 8 XCDs, each with its own L2; without the
// renumbering neighbouring tiles (which share halo lines) sit in different
// L2s.
struct Tile {
  int x, y, b;
};
__device__ inline Tile xcd_tile() {
  const int nb = gridDim.x * gridDim.y * gridDim.z;
  const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = bid & 7, s = bid >> 3;
  const int t = x * q + min(x, r) + s;
  return Tile{(int)(t % gridDim.x), (int)((t / gridDim.x) % gridDim.y),
              (int)(t / (gridDim.x * gridDim.y))};
}

// Inside flexible CG the preconditioner's launches are queued ahead of the
// host's convergence check; a workgroup of a component (batch index b) that
// has stopped returns at once.
__device__ inline bool gated_off(const FcgScal *gate, int b) { return gate && !gate[b].active; }

// ------------------------------------------------------------- MFMA conv
// out[b, y, x, co] = sum_tap sum_ci in[b, y+dy, x+dx, ci] * w[tap, co, ci].
// One workgroup per 16x16 output tile; wave w computes tile rows 4w..4w+3.
// MFMA operands: A = weights (M = 16 output channels), B = input pixels
// (N = 16 pixels of a row), K = 32 (tap, input channel) pairs.  Lane
// (n = lane & 15, g = lane >> 4) feeds A row n and B column n with the 8 K
// values of slot g; the accumulator holds D[4g + j][n], j = 0..3.
//
// UP: the input is u = BF16(a + up(bc)) (nearest upsampling of the coarser
// level bc), formed while the tile is staged; uout (optional) receives u.
//
// pout (optional, BF16 output only): also the 2x2 mean of the BF16 output,
// ((x00 + x01) + (x10 + x11)) * 0.25 (x01: the next column, x10: the next
// row), rounded to BF16, for the next level.
template <int CIN, int COUT, bool OUTF32, bool UP = false>
__global__ __launch_bounds__(256) void k_conv(const bf16 *__restrict__ in,
                                              const bf16 *__restrict__ w, void *outp, int H,
                                              int W, const bf16 *__restrict__ inc = nullptr,
                                              bf16 *__restrict__ uout = nullptr,
                                              const FcgScal *__restrict__ gate = nullptr,
                                              bf16 *__restrict__ pout = nullptr) {
  static_assert(CIN == 16 || CIN == 32, "CIN");
  static_assert(COUT % 16 == 0, "COUT");
  __shared__ __attribute__((aligned(16))) bf16 sin[(T + 2) * (T + 2) * CIN];
  __shared__ __attribute__((aligned(16))) bf16 sw[9 * COUT * CIN];
  const Tile tl = xcd_tile();
  const int b = tl.b;
  if (gated_off(gate, b)) return;
  const int x0 = tl.x * T, y0 = tl.y * T;
  const bf16 *inb = in + (size_t)b * H * W * CIN;
  constexpr int CH = CIN / 8;
  for (int e = threadIdx.x; e < (T + 2) * (T + 2) * CH; e += 256) {
    const int c = e % CH, px = e / CH;
    const int gy = wm_wrap(y0 + px / (T + 2) - 1, H), gx = wm_wrap(x0 + px % (T + 2) - 1, W);
    const size_t src = ((size_t)gy * W + gx) * CIN + c * 8;
    if (UP) {
      const bf16 *cb = inc + (size_t)b * (H / 2) * (W / 2) * CIN;
      bf16x8 va, vb, vu;
      __builtin_memcpy(&va, &inb[src], 16);
      __builtin_memcpy(&vb, &cb[((size_t)(gy >> 1) * (W / 2) + (gx >> 1)) * CIN + c * 8], 16);
#pragma unroll
      for (int k = 0; k < 8; ++k) vu[k] = (bf16)((float)va[k] + (float)vb[k]);
      __builtin_memcpy(&sin[px * CIN + c * 8], &vu, 16);
      const int ty = px / (T + 2) - 1, tx = px % (T + 2) - 1;
      if (uout && ty >= 0 && ty < T && tx >= 0 && tx < T && y0 + ty < H && x0 + tx < W)
        __builtin_memcpy(&uout[(size_t)b * H * W * CIN + src], &vu, 16);
    } else {
      *(uint4 *)&sin[px * CIN + c * 8] = *(const uint4 *)&inb[src];
    }
  }
  for (int e = threadIdx.x; e < 9 * COUT * CH; e += 256)
    *(uint4 *)&sw[e * 8] = *(const uint4 *)&w[e * 8];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = lane & 15, g = lane >> 4;
  constexpr int NCH = CIN == 16 ? 5 : 9;     // K chunks of 32
  bf16x8 zero;
#pragma unroll
  for (int k = 0; k < 8; ++k) zero[k] = (bf16)0.0f;
#pragma unroll
  for (int mb = 0; mb < COUT / 16; ++mb) {
    f32x4 acc[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) acc[rr] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int tap = CIN == 16 ? 2 * c + (g >> 1) : c;
      const int cin0 = CIN == 16 ? 8 * (g & 1) : 8 * g;
      const bool live = tap < 9;
      const int tp = live ? tap : 8;
      const int dy = tp / 3 - 1, dx = tp % 3 - 1;
      const bf16x8 A = live ? *(const bf16x8 *)&sw[(tp * COUT + mb * 16 + n) * CIN + cin0] : zero;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int ty = wave * 4 + rr;
        const bf16x8 Bv =
            live ? *(const bf16x8 *)&sin[((ty + dy + 1) * (T + 2) + (n + dx + 1)) * CIN + cin0] : zero;
        acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bv, acc[rr], 0, 0, 0);
      }
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int gy = y0 + wave * 4 + rr, gx = x0 + n;
      if (gy < H && gx < W) {
        const size_t o = (((size_t)b * H + gy) * W + gx) * COUT + mb * 16 + 4 * g;
        if (OUTF32) {
          float *out = (float *)outp;
          *(float4 *)&out[o] = make_float4(acc[rr][0], acc[rr][1], acc[rr][2], acc[rr][3]);
        } else {
          bf16 *out = (bf16 *)outp;
#pragma unroll
          for (int j = 0; j < 4; ++j) out[o + j] = (bf16)acc[rr][j];
        }
      }
    }
    if (!OUTF32 && pout) {
      // rows (rr, rr + 1) of this lane, columns (n, n ^ 1) of lanes (lane, lane ^ 1)
      const int Ho = H / 2, Wo = W / 2;
#pragma unroll
      for (int rr = 0; rr < 4; rr += 2) {
        const int Y = (y0 + wave * 4 + rr) >> 1, X = (x0 + n) >> 1;
        const bool st = (n & 1) == 0 && Y < Ho && X < Wo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float a00 = (float)(bf16)acc[rr][j], a10 = (float)(bf16)acc[rr + 1][j];
          const float a01 = __shfl_xor(a00, 1, 64), a11 = __shfl_xor(a10, 1, 64);
          if (st)
            pout[(((size_t)b * Ho + Y) * Wo + X) * COUT + mb * 16 + 4 * g + j] =
                (bf16)(((a00 + a01) + (a10 + a11)) * 0.25f);
        }
      }
    }
  }
}

// Fixed-order workgroup sum of one double; result in thread 0.
__device__ inline double block_sum1(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// ------------------------------------------------------- fine-level kernels
// x0 tile with a halo of HALO pixels, (float)(r * inv_s), periodic.  Staged
// in two passes, every thread's loads issued before any is used, so the
// tile costs one memory latency.
template <int HALO>
struct X0Tile {
  static constexpr int E = T + 2 * HALO, N = E * E, NIT = (N + 255) / 256;
  double v[NIT];
};
template <int HALO>
__device__ inline void x0_issue(X0Tile<HALO> &X, const double *__restrict__ r, int x0, int y0,
                                int H, int W) {
  constexpr int E = X0Tile<HALO>::E;
#pragma unroll
  for (int i = 0; i < X0Tile<HALO>::NIT; ++i) {
    const int e = threadIdx.x + 256 * i;
    X.v[i] = 0.0;
    if (e < X0Tile<HALO>::N) {
      const int gy = wm_wrap(y0 + e / E - HALO, H), gx = wm_wrap(x0 + e % E - HALO, W);
      X.v[i] = r[(size_t)gy * W + gx];
    }
  }
}
template <int HALO>
__device__ inline void x0_put(float *xs, const X0Tile<HALO> &X, double is) {
#pragma unroll
  for (int i = 0; i < X0Tile<HALO>::NIT; ++i) {
    const int e = threadIdx.x + 256 * i;
    if (e < X0Tile<HALO>::N) xs[e] = (float)(X.v[i] * is);
  }
}
template <int HALO>
__device__ inline void load_x0(float *xs, const double *__restrict__ r, double is, int x0,
                               int y0, int H, int W) {
  X0Tile<HALO> X;
  x0_issue<HALO>(X, r, x0, y0, H, W);
  x0_put<HALO>(xs, X, is);
}

// a0 (C0 channels) at tile-local pixel (py, px) of an x0 tile of edge E whose
// origin sits `o` pixels before the pixel's own frame.
template <int E>
__device__ inline void lift_at(float *a, const float *xs, const float *__restrict__ wl, int py,
                               int px) {
#pragma unroll
  for (int co = 0; co < C0; ++co) a[co] = 0.0f;
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const float v = xs[(py + tap / 3 - 1) * E + (px + tap % 3 - 1)];
#pragma unroll
    for (int co = 0; co < C0; ++co) a[co] = fmaf(v, wl[tap * C0 + co], a[co]);
  }
}

// Folded fine-level weights (k_fold), FP32:
//   FK[5][5]            lift then proj as one 5x5 stencil on x0
//   FP[py][px][a][b][c] proj applied to up(b1): fine pixels of parity
//                       (py, px) see coarse pixels at offsets a - 1 + py,
//                       b - 1 + px (a, b in {0, 1})
//   FL[r][c][co]        pool(lift(x0)) as one 4x4 stencil per channel on the
//                       x0 rows 2Y-1..2Y+2, columns 2X-1..2X+2
constexpr int FK_OFF = 0, FP_OFF = 32, FL_OFF = 32 + 256, NFOLD = 32 + 256 + 256;

__global__ void k_fold(const float *__restrict__ wl, const float *__restrict__ wp,
                       float *__restrict__ f) {
  const int t = threadIdx.x;
  if (t < 25) {
    const int ky = t / 5 - 2, kx = t % 5 - 2;
    double v = 0.0;
    for (int tp = 0; tp < 9; ++tp)
      for (int tl = 0; tl < 9; ++tl)
        if (tp / 3 - 1 + tl / 3 - 1 == ky && tp % 3 - 1 + tl % 3 - 1 == kx)
          for (int ci = 0; ci < C0; ++ci) v += (double)wp[tp * C0 + ci] * wl[tl * C0 + ci];
    f[FK_OFF + t] = (float)v;
  }
  for (int e = t; e < 256; e += blockDim.x) {
    // FP: e = (((py * 2 + px) * 2 + a) * 2 + b) * C0 + ci
    const int ci = e % C0, bb = (e / C0) % 2, a = (e / (2 * C0)) % 2, px = (e / (4 * C0)) % 2,
              py = e / (8 * C0);
    float v = 0.0f;
    for (int tp = 0; tp < 9; ++tp) {
      const int dy = tp / 3 - 1, dx = tp % 3 - 1;
      const int cy = (py + dy) >> 1, cx = (px + dx) >> 1;   // floor
      if (cy == a - 1 + py && cx == bb - 1 + px) v += wp[tp * C0 + ci];
    }
    f[FP_OFF + e] = v;
    // FL: e = (r * 4 + c) * C0 + co
    const int co = e % C0, c = (e / C0) % 4, r = e / (4 * C0);
    double w = 0.0;
    for (int chy = 0; chy < 2; ++chy)
      for (int chx = 0; chx < 2; ++chx)
        for (int tl = 0; tl < 9; ++tl)
          if (chy + tl / 3 == r && chx + tl % 3 == c) w += wl[tl * C0 + co];
    f[FL_OFF + e] = (float)(0.25 * w);
  }
}

// d1 = pool(lift(r / s)) with the folded 4x4 stencils, one 16x16 fine tile
// -> 8x8 level-1 pixels.
__global__ __launch_bounds__(256) void k_lift_pool(const double *__restrict__ r, size_t rstride,
                                                   const double *__restrict__ sc,
                                                   const float *__restrict__ fold,
                                                   bf16 *__restrict__ d1, int H, int W, int B,
                                                   const FcgScal *__restrict__ gate) {
  __shared__ float xs[(T + 2) * (T + 2)];
  __shared__ __attribute__((aligned(16))) float fl[16 * C0];
  const Tile tl = xcd_tile();
  const int b = __builtin_amdgcn_readfirstlane(tl.b);
  if (gated_off(gate, b)) return;
  const int x0 = tl.x * T, y0 = tl.y * T;
  static_assert(16 * C0 == 256, "one weight per thread");
  X0Tile<1> xt;
  x0_issue<1>(xt, r + (size_t)b * rstride, x0, y0, H, W);
  const float flv = fold[FL_OFF + threadIdx.x];
  x0_put<1>(xs, xt, sc[B + b]);
  fl[threadIdx.x] = flv;
  __syncthreads();
  const int H1 = H / 2, W1 = W / 2;
  // thread: channels 4 c4 .. 4 c4 + 3 of level-1 pixel q (64 pixels x 4)
  static_assert((T / 2) * (T / 2) * (C0 / 4) == 256, "one pass");
  const int c4 = threadIdx.x % (C0 / 4), q = threadIdx.x / (C0 / 4);
  const int qy = q / (T / 2), qx = q % (T / 2);
  const int Y = y0 / 2 + qy, X = x0 / 2 + qx;
  if (Y >= H1 || X >= W1) return;
  float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int rr = 0; rr < 4; ++rr)
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const float xv = xs[(2 * qy + rr) * (T + 2) + 2 * qx + cc];
      const float4 w = *(const float4 *)&fl[(rr * 4 + cc) * C0 + 4 * c4];
      v[0] = fmaf(w.x, xv, v[0]);
      v[1] = fmaf(w.y, xv, v[1]);
      v[2] = fmaf(w.z, xv, v[2]);
      v[3] = fmaf(w.w, xv, v[3]);
    }
  typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
  bf16x4 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = (bf16)v[k];
  __builtin_memcpy(&d1[(((size_t)b * H1 + Y) * W1 + X) * C0 + 4 * c4], &o, 8);
}

// u0 = lift(x0) + up(b1) on the tile plus a 1-pixel halo, in LDS.
__device__ inline void build_u0(float *us, const float *xs, const float *__restrict__ wl,
                                const bf16 *__restrict__ b1b, int x0, int y0, int H, int W) {
  const int W1 = W / 2;
  for (int e = threadIdx.x; e < (T + 2) * (T + 2); e += 256) {
    const int uy = e / (T + 2), ux = e % (T + 2);
    float a[C0];
    lift_at<T + 4>(a, xs, wl, uy + 1, ux + 1);
    const int gy = pmod(y0 + uy - 1, H), gx = pmod(x0 + ux - 1, W);
    const bf16 *bp = b1b + ((size_t)(gy / 2) * W1 + gx / 2) * C0;
#pragma unroll
    for (int co = 0; co < C0; ++co) us[e * C0 + co] = a[co] + (float)bp[co];
  }
}

// z = s * proj(lift(r / s) + up(b1)) with the folded weights:
// y = FK * x0 + sum over the 2x2 coarse neighbours of FP[parity] . b1;
// y itself is stored as well when y != null.
//
// With dots != null (inside flexible CG) the kernel also leaves this tile's
// partial sums of r.z and z.q in dots[(b * 2 + k) * tiles + tile], k = 0, 1
// (z.q = 0 on the component's first iteration), fixed-order tree sums.
__global__ __launch_bounds__(256) void k_up_proj(const double *__restrict__ r, size_t rstride,
                                                 const double *__restrict__ sc,
                                                 const float *__restrict__ fold,
                                                 const bf16 *__restrict__ b1,
                                                 double *__restrict__ z, size_t zstride,
                                                 float *__restrict__ y, int H, int W, int B,
                                                 const double *__restrict__ q,
                                                 double *__restrict__ dots,
                                                 const FcgScal *__restrict__ S) {
  constexpr int CE = T / 2 + 2;                    // coarse tile edge (halo 1)
  // coarse pixels 20 floats apart (16 channels + 4): the lanes' 16-byte reads
  // of neighbouring pixels fall in different LDS banks
  constexpr int CS = C0 + 4;
  __shared__ float xs[(T + 4) * (T + 4)];
  __shared__ __attribute__((aligned(16))) float bs[CE * CE * CS];
  // parity groups 68 floats apart, so a wave's four parities use different banks
  constexpr int PS = 4 * C0 + 4;
  __shared__ __attribute__((aligned(16))) float fp[4 * PS];
  __shared__ float fk[25];
  __shared__ double sh[8];
  const Tile tl = xcd_tile();
  const int b = __builtin_amdgcn_readfirstlane(tl.b);
  if (dots && gated_off(S, b)) return;
  const int x0 = tl.x * T, y0 = tl.y * T;
  const int ty = threadIdx.x / T, tx = threadIdx.x % T;
  const int gy = y0 + ty, gx = x0 + tx;
  const bool in = gy < H && gx < W;
  const size_t p = (size_t)gy * W + gx;
  // every load of the tile goes out before any is used: one memory latency
  const bool first = dots && S[b].it == 0;
  const double rv = (dots && in) ? r[(size_t)b * rstride + p] : 0.0;
  const double qv = (dots && in && !first) ? q[(size_t)b * rstride + p] : 0.0;
  X0Tile<2> xt;
  x0_issue<2>(xt, r + (size_t)b * rstride, x0, y0, H, W);
  const int H1 = H / 2, W1 = W / 2;
  const bf16 *b1b = b1 + (size_t)b * H1 * W1 * C0;
  static_assert(CE * CE * (C0 / 8) <= 256, "one coarse load per thread");
  const int c8 = threadIdx.x % (C0 / 8), cq = threadIdx.x / (C0 / 8);
  const bool cin = threadIdx.x < CE * CE * (C0 / 8);
  bf16x8 cv;
  if (cin) {
    const int Y = wm_wrap(y0 / 2 + cq / CE - 1, H1), X = wm_wrap(x0 / 2 + cq % CE - 1, W1);
    __builtin_memcpy(&cv, &b1b[((size_t)Y * W1 + X) * C0 + 8 * c8], 16);
  }
  const float fpv = fold[FP_OFF + threadIdx.x];
  const float fkv = threadIdx.x < 25 ? fold[FK_OFF + threadIdx.x] : 0.0f;
  const double scale = sc[b];
  x0_put<2>(xs, xt, sc[B + b]);
  if (cin) {
    float4 *dst = (float4 *)&bs[cq * CS + 8 * c8];
    dst[0] = make_float4((float)cv[0], (float)cv[1], (float)cv[2], (float)cv[3]);
    dst[1] = make_float4((float)cv[4], (float)cv[5], (float)cv[6], (float)cv[7]);
  }
  fp[(threadIdx.x / (4 * C0)) * PS + threadIdx.x % (4 * C0)] = fpv;
  if (threadIdx.x < 25) fk[threadIdx.x] = fkv;
  __syncthreads();
  float v = 0.0f;
#pragma unroll
  for (int k = 0; k < 25; ++k) v = fmaf(fk[k], xs[(ty + k / 5) * (T + 4) + tx + k % 5], v);
  const int py = ty & 1, px = tx & 1;
  const float *fpp = fp + (py * 2 + px) * PS;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      // coarse pixel (ty>>1) + a - 1 + py in the tile frame, +1 for the halo;
      // 16 channels as four 16-byte LDS reads of features and of weights
      const float4 *bp = (const float4 *)(bs + ((ty / 2 + a + py) * CE + (tx / 2 + bb + px)) * CS);
      const float4 *wv = (const float4 *)(fpp + (a * 2 + bb) * C0);
#pragma unroll
      for (int c4 = 0; c4 < C0 / 4; ++c4) {
        const float4 f = bp[c4], w = wv[c4];
        v = fmaf(w.x, f.x, v);
        v = fmaf(w.y, f.y, v);
        v = fmaf(w.z, f.z, v);
        v = fmaf(w.w, f.w, v);
      }
    }
  const double zv = (double)v * scale;
  if (in) {
    z[(size_t)b * zstride + p] = zv;
    if (y) y[(size_t)b * H * W + p] = v;
  }
  if (dots) {
    const double zr = in ? zv * rv : 0.0;
    const double zq = (in && !first) ? zv * qv : 0.0;
    const int tiles = gridDim.x * gridDim.y, tile = tl.y * gridDim.x + tl.x;
    // fixed order: a butterfly within each wave, then the 4 wave totals
    const double w0 = wm_wave_sum(zr), w1 = wm_wave_sum(zq);
    if ((threadIdx.x & 63) == 0) {
      sh[threadIdx.x >> 6] = w0;
      sh[4 + (threadIdx.x >> 6)] = w1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      dots[(size_t)(b * 2 + 0) * tiles + tile] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
      dots[(size_t)(b * 2 + 1) * tiles + tile] = ((sh[4] + sh[5]) + sh[6]) + sh[7];
    }
  }
}

// ---------------------------------------------------------------- backward
// Adjoint of up: out[q] = BF16((g00 + g01) + (g10 + g11)).
__global__ void k_up_t(const float *__restrict__ g, bf16 *__restrict__ out, int B, int H, int W,
                       int C) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t n = (size_t)B * Ho * Wo * C;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = e % C;
    const size_t q = e / C;
    const int X = q % Wo, Y = (q / Wo) % Ho, b = q / ((size_t)Wo * Ho);
    const float *p = g + (((size_t)b * H + 2 * Y) * W + 2 * X) * C + c;
    out[e] = (bf16)((p[0] + p[C]) + (p[(size_t)W * C] + p[(size_t)W * C + C]));
  }
}

// out = BF16(du + 0.25 * up(dd)): gradient reaching a level's activation
// through the up+add path (du) and through the pool feeding the next level.
__global__ void k_add_unpool(const float *__restrict__ du, const float *__restrict__ dd,
                             bf16 *__restrict__ out, int B, int H, int W, int C) {
  const size_t n = (size_t)B * H * W * C;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = e % C;
    const size_t q = e / C;
    const int X = q % W, Y = (q / W) % H, b = q / ((size_t)W * H);
    const float v = du[e] + 0.25f * dd[(((size_t)b * (H / 2) + Y / 2) * (W / 2) + X / 2) * C + c];
    out[e] = (bf16)v;
  }
}

// Backward of y = proj(u0) for the loss mean((y - e/s)^2):
//   dy = 2 (y - e/s) / N, dW_proj = sum u0(p+off) dy(p), du0 = proj^T dy,
// plus the loss.  Workgroup k handles tiles k, k + grid, ... in order and
// leaves one partial row (and one loss partial).
__global__ __launch_bounds__(256) void k_proj_bwd(
    const double *__restrict__ r, size_t rstride, const double *__restrict__ e, size_t estride,
    const double *__restrict__ sc, const float *__restrict__ y, const float *__restrict__ wl,
    const float *__restrict__ wp, const bf16 *__restrict__ b1, float *__restrict__ du0,
    float *__restrict__ part, double *__restrict__ lossp, double gs, int H, int W, int B,
    int ntx, int nty, int ntiles) {
  __shared__ float xs[(T + 4) * (T + 4)];
  __shared__ float us[(T + 2) * (T + 2) * C0];
  __shared__ float dys[(T + 2) * (T + 2)];
  __shared__ double sh[256];
  const int ty = threadIdx.x / T, tx = threadIdx.x % T;
  double l2 = 0.0;
  float acc = 0.0f;                                  // dW_proj[threadIdx.x] (< 9*C0)
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int b = t / (ntx * nty), tt = t % (ntx * nty);
    const int x0 = (tt % ntx) * T, y0 = (tt / ntx) * T;
    const double is = sc[B + b];
    __syncthreads();
    load_x0<2>(xs, r + (size_t)b * rstride, is, x0, y0, H, W);
    for (int q = threadIdx.x; q < (T + 2) * (T + 2); q += 256) {
      const int gy = pmod(y0 + q / (T + 2) - 1, H), gx = pmod(x0 + q % (T + 2) - 1, W);
      const size_t p = (size_t)gy * W + gx;
      const double d = (double)y[(size_t)b * H * W + p] - e[(size_t)b * estride + p] * is;
      dys[q] = (float)(gs * d);
    }
    __syncthreads();
    build_u0(us, xs, wl, b1 + (size_t)b * (H / 2) * (W / 2) * C0, x0, y0, H, W);
    __syncthreads();
    if (y0 + ty < H && x0 + tx < W) {
      const size_t p = (size_t)(y0 + ty) * W + (x0 + tx);
      const double d = (double)y[(size_t)b * H * W + p] - e[(size_t)b * estride + p] * is;
      l2 += d * d;
      float g[C0];
#pragma unroll
      for (int ci = 0; ci < C0; ++ci) g[ci] = 0.0f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        // dy at p - off(tap)
        const float v = dys[(ty + 1 - (tap / 3 - 1)) * (T + 2) + (tx + 1 - (tap % 3 - 1))];
#pragma unroll
        for (int ci = 0; ci < C0; ++ci) g[ci] = fmaf(v, wp[tap * C0 + ci], g[ci]);
      }
      float *o = du0 + (((size_t)b * H + y0 + ty) * W + x0 + tx) * C0;
#pragma unroll
      for (int ci = 0; ci < C0; ci += 4) *(float4 *)&o[ci] = make_float4(g[ci], g[ci + 1], g[ci + 2], g[ci + 3]);
    }
    if (threadIdx.x < 9 * C0) {
      const int tap = threadIdx.x / C0, ci = threadIdx.x % C0;
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      for (int q = 0; q < T * T; ++q) {
        const int qy = q / T, qx = q % T;
        if (y0 + qy >= H || x0 + qx >= W) continue;
        acc = fmaf(us[((qy + 1 + dy) * (T + 2) + (qx + 1 + dx)) * C0 + ci],
                   dys[(qy + 1) * (T + 2) + qx + 1], acc);
      }
    }
  }
  __syncthreads();
  const double tot = block_sum1(l2, sh);
  if (threadIdx.x == 0) lossp[blockIdx.x] = tot;
  if (threadIdx.x < 9 * C0) part[(size_t)blockIdx.x * (9 * C0) + threadIdx.x] = acc;
}

// dW_lift[tap][co] = sum_p x0(p+off) * da0(p, co), da0 = du0 + 0.25 up(dd1);
// tiles as in k_proj_bwd.
__global__ __launch_bounds__(256) void k_lift_bwd(const double *__restrict__ r, size_t rstride,
                                                  const double *__restrict__ sc,
                                                  const float *__restrict__ du0,
                                                  const float *__restrict__ dd1,
                                                  float *__restrict__ part, int H, int W, int B,
                                                  int ntx, int nty, int ntiles) {
  __shared__ float xs[(T + 2) * (T + 2)];
  __shared__ float das[T * T * C0];
  const int W1 = W / 2;
  float acc = 0.0f;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int b = t / (ntx * nty), tt = t % (ntx * nty);
    const int x0 = (tt % ntx) * T, y0 = (tt / ntx) * T;
    __syncthreads();
    load_x0<1>(xs, r + (size_t)b * rstride, sc[B + b], x0, y0, H, W);
    for (int e = threadIdx.x; e < T * T * C0; e += 256) {
      const int co = e % C0, q = e / C0;
      const int gy = y0 + q / T, gx = x0 + q % T;
      float v = 0.0f;
      if (gy < H && gx < W)
        v = du0[(((size_t)b * H + gy) * W + gx) * C0 + co] +
            0.25f * dd1[(((size_t)b * (H / 2) + gy / 2) * W1 + gx / 2) * C0 + co];
      das[e] = v;
    }
    __syncthreads();
    if (threadIdx.x < 9 * C0) {
      const int tap = threadIdx.x / C0, co = threadIdx.x % C0;
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      for (int q = 0; q < T * T; ++q) {
        const int qy = q / T, qx = q % T;
        acc = fmaf(xs[(qy + 1 + dy) * (T + 2) + qx + 1 + dx], das[q * C0 + co], acc);
      }
    }
  }
  if (threadIdx.x < 9 * C0) part[(size_t)blockIdx.x * (9 * C0) + threadIdx.x] = acc;
}

// Weight gradient of an MFMA layer: dW[tap][co][ci] = sum_p X(p+off, ci) G(p, co),
// BF16 inputs, FP32 sums.  Workgroup k sums tiles k, k + grid, ... in order.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_dw(const bf16 *__restrict__ X,
                                                 const bf16 *__restrict__ G,
                                                 float *__restrict__ part, int H, int W,
                                                 int ntx, int nty, int ntiles) {
  __shared__ float xs[(T + 2) * (T + 2) * CIN];
  __shared__ bf16 gsh[T * T * COUT];
  constexpr int NO = 9 * CIN * COUT;
  constexpr int PER = (NO + 255) / 256;
  float acc[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) acc[k] = 0.0f;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int b = t / (ntx * nty), tt = t % (ntx * nty);
    const int x0 = (tt % ntx) * T, y0 = (tt / ntx) * T;
    __syncthreads();
    for (int e = threadIdx.x; e < (T + 2) * (T + 2) * CIN; e += 256) {
      const int c = e % CIN, px = e / CIN;
      const int gy = pmod(y0 + px / (T + 2) - 1, H), gx = pmod(x0 + px % (T + 2) - 1, W);
      xs[e] = (float)X[(((size_t)b * H + gy) * W + gx) * CIN + c];
    }
    for (int e = threadIdx.x; e < T * T * COUT; e += 256) {
      const int c = e % COUT, q = e / COUT;
      const int gy = y0 + q / T, gx = x0 + q % T;
      gsh[e] = (gy < H && gx < W) ? G[(((size_t)b * H + gy) * W + gx) * COUT + c] : (bf16)0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int o = threadIdx.x + 256 * k;
      if (o >= NO) break;
      const int ci = o % CIN, co = (o / CIN) % COUT, tap = o / (CIN * COUT);
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      float a = acc[k];
      for (int q = 0; q < T * T; ++q) {
        const int qy = q / T, qx = q % T;
        a = fmaf(xs[((qy + 1 + dy) * (T + 2) + qx + 1 + dx) * CIN + ci], (float)gsh[q * COUT + co], a);
      }
      acc[k] = a;
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int o = threadIdx.x + 256 * k;
    if (o < NO) part[(size_t)blockIdx.x * NO + o] = acc[k];
  }
}

// grad[o] = sum over nblk partial rows: 32 outputs per workgroup, eight
// row groups (rows g, g + 8, ...) added in a fixed order.
__global__ __launch_bounds__(256) void k_reduce_rows(const float *__restrict__ part, int nblk,
                                                     int n, float *__restrict__ grad) {
  __shared__ float sh[8][32];
  const int ol = threadIdx.x % 32, g = threadIdx.x / 32;
  const int o = blockIdx.x * 32 + ol;
  float s = 0.0f;
  if (o < n)
    for (int k = g; k < nblk; k += 8) s += part[(size_t)k * n + o];
  sh[g][ol] = s;
  __syncthreads();
  if (g == 0 && o < n) {
    float t = sh[0][ol];
#pragma unroll
    for (int k = 1; k < 8; ++k) t += sh[k][ol];
    grad[o] = t;
  }
}

__global__ void k_sum_loss(const double *__restrict__ lossp, int nblk, double inv_n,
                           double *__restrict__ out) {
  __shared__ double sh[256];
  double v = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) v += lossp[k];
  const double t = block_sum1(v, sh);
  if (threadIdx.x == 0) *out = t * inv_n;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; c1 = 1/(1 - b1^t),
// c2 = 1/(1 - b2^t).
__global__ void k_adam(float *__restrict__ th, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const float gg = g[o];
  const float mm = b1 * m[o] + (1.0f - b1) * gg;
  const float vv = b2 * v[o] + (1.0f - b2) * gg * gg;
  m[o] = mm;
  v[o] = vv;
  th[o] = th[o] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// BF16 copies of the MFMA layers: wb in the [tap][cout][cin] layout, wt the
// adjoint's weights wt[tap'][ci][co] = w[8 - tap'][co][ci].
__global__ void k_weights_bf16(const float *__restrict__ th, bf16 *__restrict__ wb,
                               bf16 *__restrict__ wt) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= NPARAM) return;
  int l = 0;
  while (l + 1 < NLAYER && o >= loff(l + 1)) ++l;
  if (l == L_LIFT || l == L_PROJ) return;
  const int ci_n = LCIN[l], co_n = LCOUT[l];
  const int k = o - loff(l);
  const int ci = k % ci_n, co = (k / ci_n) % co_n, tap = k / (ci_n * co_n);
  const bf16 v = (bf16)th[o];
  wb[o] = v;
  wt[loff(l) + ((8 - tap) * ci_n + ci) * co_n + co] = v;
}

// Pair snapshot of the plain solver (iteration k): r_k and x_k of the three
// components, and their scales s = |r_k| / sqrt(n) from the solver's own sum.
__global__ void k_snap(const double *__restrict__ r, const double *__restrict__ df, int nx,
                       int ny, int gx, size_t ps, const CgScal *__restrict__ S,
                       double *__restrict__ sr, double *__restrict__ sx, double *__restrict__ ssc) {
  const size_t n = (size_t)nx * ny;
  const int l = blockIdx.y;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const int ii = t % nx, jj = t / nx;
    sr[l * n + t] = r[l * n + t];
    sx[l * n + t] = df[l * ps + (size_t)(ii + 2) + (size_t)gx * (jj + 2)];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = sqrt(S[l].sumr / (double)n);
    if (!(s > 0.0)) s = 1.0;
    ssc[l] = s;
    ssc[3 + l] = 1.0 / s;
  }
}

// e = x_final - x_k, in place over the snapshot's x.
__global__ void k_snap_fin(const double *__restrict__ df, int nx, int ny, int gx, size_t ps,
                           double *__restrict__ sx) {
  const size_t n = (size_t)nx * ny;
  const int l = blockIdx.y;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const int ii = t % nx, jj = t / nx;
    sx[l * n + t] = df[l * ps + (size_t)(ii + 2) + (size_t)gx * (jj + 2)] - sx[l * n + t];
  }
}

// Scales of three host-supplied images (rms over the image, fixed order).
__global__ void k_img_scale(const double *__restrict__ r, size_t n, double *__restrict__ sc) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  double v = 0.0;
  for (size_t t = threadIdx.x; t < n; t += 256) v += r[b * n + t] * r[b * n + t];
  const double tot = block_sum1(v, sh);
  if (threadIdx.x == 0) {
    double s = sqrt(tot / (double)n);
    if (!(s > 0.0)) s = 1.0;
    sc[b] = s;
    sc[3 + b] = 1.0 / s;
  }
}

}  // namespace

// ------------------------------------------------------------------ state
struct wm_pc {
  int H = 0, W = 0;
  static constexpr int B = 3;
  float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  bf16 *wb = nullptr, *wt = nullptr;
  float *fold = nullptr;         // folded fine-level weights (k_fold)
  int adam_t = 0;
  bool trained = false;
  // activations (BF16, NHWC)
  bf16 *d1 = nullptr, *a1 = nullptr, *d2 = nullptr, *a2 = nullptr, *d3 = nullptr, *a3 = nullptr;
  bf16 *u2 = nullptr, *b2 = nullptr, *u1 = nullptr, *b1 = nullptr;
  // training
  float *y = nullptr, *du0 = nullptr, *du1 = nullptr, *du2 = nullptr;
  float *dd1 = nullptr, *dd2 = nullptr, *dd3 = nullptr;
  bf16 *db1 = nullptr, *db2 = nullptr, *da3 = nullptr, *da2 = nullptr, *da1 = nullptr;
  double *zt = nullptr;          // training forward output
  float *part = nullptr;         // weight-gradient partial rows
  size_t part_elems = 0;
  double *lossp = nullptr, *losses = nullptr;
  int nlossp = 0, losses_cap = 0;
  double *isc = nullptr;         // scales of host-supplied images
  double *himg = nullptr;        // host-supplied images [2][3][n]
  // pairs recorded from the plain solver
  int nslot = 0, per_solve = 0, filled = 0, next = 0;
  double *pr = nullptr, *pe = nullptr, *psc = nullptr;
  int snap_slot[64];
  int nsnap = 0;
  int last_fcg[3] = {0, 0, 0}, last_fallback[3] = {0, 0, 0};
};

namespace {

template <class Tp>
int dalloc(wm_ctx *ctx, Tp **p, size_t n) {
  WM_HIP(hipMalloc((void **)p, sizeof(Tp) * (n ? n : 1)));
  return WM_OK;
}

size_t lvl(const wm_pc *pc, int l, int c) {
  return (size_t)wm_pc::B * (pc->H >> l) * (pc->W >> l) * c;
}

dim3 tiles(int H, int W, int B) { return dim3(wm_cdiv(W, T), wm_cdiv(H, T), B); }

int ew_blocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

template <int CIN, int COUT, bool F32>
int conv(wm_ctx *ctx, const bf16 *in, const bf16 *w, void *out, int H, int W,
         const FcgScal *gate = nullptr, bf16 *pool = nullptr) {
  hipLaunchKernelGGL((k_conv<CIN, COUT, F32>), tiles(H, W, wm_pc::B), dim3(256), 0, ctx->stream,
                     in, w, out, H, W, nullptr, nullptr, gate, pool);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

template <int CIN, int COUT>
int conv_up(wm_ctx *ctx, const bf16 *a, const bf16 *bc, bf16 *uout, const bf16 *w, bf16 *out,
            int H, int W, const FcgScal *gate) {
  hipLaunchKernelGGL((k_conv<CIN, COUT, false, true>), tiles(H, W, wm_pc::B), dim3(256), 0,
                     ctx->stream, a, w, (void *)out, H, W, bc, uout, gate);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// Weight gradient of layer l into grad (partials then fixed-order rows).
template <int CIN, int COUT>
int conv_dw(wm_ctx *ctx, wm_pc *pc, int l, const bf16 *X, const bf16 *G, int H, int W) {
  const int ntx = wm_cdiv(W, T), nty = wm_cdiv(H, T), ntiles = ntx * nty * wm_pc::B;
  const int nblk = std::min(ntiles, 512);
  constexpr int NO = 9 * CIN * COUT;
  hipLaunchKernelGGL((k_conv_dw<CIN, COUT>), dim3(nblk), dim3(256), 0, ctx->stream, X, G, pc->part,
                     H, W, ntx, nty, ntiles);
  hipLaunchKernelGGL(k_reduce_rows, dim3(wm_cdiv(NO, 32)), dim3(256), 0, ctx->stream, pc->part,
                     nblk, NO, pc->grad + loff(l));
  WM_HIP(hipGetLastError());
  return WM_OK;
}

// The activations of one forward pass (BF16, NHWC, the three images batched).
struct Acts {
  bf16 *d1, *a1, *d2, *a2, *d3, *a3, *u2, *b2, *u1, *b1;
};
Acts acts_of(const wm_pc *pc) {
  return Acts{pc->d1, pc->a1, pc->d2, pc->a2, pc->d3, pc->a3, pc->u2, pc->b2, pc->u1, pc->b1};
}

// Forward pass over three H x W images r (stride rstride) with scales sc
// (device, s[3] then 1/s[3]), activations in X; z gets s * y (stride
// zstride), y (optional) the net output.
int forward_on(wm_ctx *ctx, const wm_pc *pc, const Acts &X, int H, int W, const double *r,
               size_t rstride, const double *sc, double *z, size_t zstride, float *y,
               const double *q = nullptr, double *dots = nullptr, const FcgScal *S = nullptr) {
  const int B = wm_pc::B;
  const bf16 *wb = pc->wb;
  hipStream_t st = ctx->stream;
  const FcgScal *gate = dots ? S : nullptr;
  hipLaunchKernelGGL(k_lift_pool, tiles(H, W, B), dim3(256), 0, st, r, rstride, sc, pc->fold,
                     X.d1, H, W, B, gate);
  // d2 = pool(a1) and d3 = pool(a2) leave the D1 and D2 convolutions with a1, a2
  WM_TRY((conv<C0, C1, false>(ctx, X.d1, wb + loff(L_D1), X.a1, H >> 1, W >> 1, gate, X.d2)));
  WM_TRY((conv<C1, C2, false>(ctx, X.d2, wb + loff(L_D2), X.a2, H >> 2, W >> 2, gate, X.d3)));
  WM_TRY((conv<C2, C2, false>(ctx, X.d3, wb + loff(L_D3), X.a3, H >> 3, W >> 3, gate)));
  // u2 = a2 + up(a3) and u1 = a1 + up(b2) are formed inside the U2/U1
  // convolutions; training keeps them for the weight gradients
  const bool keep = y != nullptr;
  WM_TRY((conv_up<C2, C1>(ctx, X.a2, X.a3, keep ? X.u2 : nullptr, wb + loff(L_U2), X.b2, H >> 2,
                          W >> 2, gate)));
  WM_TRY((conv_up<C1, C0>(ctx, X.a1, X.b2, keep ? X.u1 : nullptr, wb + loff(L_U1), X.b1, H >> 1,
                          W >> 1, gate)));
  hipLaunchKernelGGL(k_up_proj, tiles(H, W, B), dim3(256), 0, st, r, rstride, sc, pc->fold, X.b1,
                     z, zstride, y, H, W, B, q, dots, S);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
int forward(wm_ctx *ctx, wm_pc *pc, const double *r, size_t rstride, const double *sc, double *z,
            size_t zstride, float *y, const double *q = nullptr, double *dots = nullptr,
            const FcgScal *S = nullptr) {
  return forward_on(ctx, pc, acts_of(pc), pc->H, pc->W, r, rstride, sc, z, zstride, y, q, dots, S);
}

// Loss and gradient (into pc->grad) of the pair (r, e) with scales sc;
// the loss lands in *loss_dev.
int backward(wm_ctx *ctx, wm_pc *pc, const double *r, const double *e, size_t stride,
             const double *sc, double *loss_dev) {
  const int H = pc->H, W = pc->W, B = wm_pc::B;
  const float *th = pc->theta;
  const bf16 *wt = pc->wt;
  hipStream_t st = ctx->stream;
  WM_TRY(forward(ctx, pc, r, stride, sc, pc->zt, (size_t)H * W, pc->y));
  const double ntot = (double)B * H * W;
  const dim3 t0 = tiles(H, W, B);
  const int ntiles0 = (int)(t0.x * t0.y * t0.z);
  const int nb0 = std::min(ntiles0, FINE_BLOCKS);
  hipLaunchKernelGGL(k_proj_bwd, dim3(nb0), dim3(256), 0, st, r, stride, e, stride, sc, pc->y,
                     th + loff(L_LIFT), th + loff(L_PROJ), pc->b1, pc->du0, pc->part, pc->lossp,
                     2.0 / ntot, H, W, B, (int)t0.x, (int)t0.y, ntiles0);
  hipLaunchKernelGGL(k_reduce_rows, dim3(wm_cdiv(9 * C0, 32)), dim3(256), 0, st, pc->part, nb0,
                     9 * C0, pc->grad + loff(L_PROJ));
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, st, pc->lossp, nb0, 1.0 / ntot, loss_dev);
  WM_HIP(hipGetLastError());
  // U1: b1 = U1(u1); u0 = a0 + up(b1)
  hipLaunchKernelGGL(k_up_t, dim3(ew_blocks(lvl(pc, 1, C0))), dim3(256), 0, st, pc->du0, pc->db1,
                     B, H, W, C0);
  WM_TRY((conv_dw<C1, C0>(ctx, pc, L_U1, pc->u1, pc->db1, H >> 1, W >> 1)));
  WM_TRY((conv<C0, C1, true>(ctx, pc->db1, wt + loff(L_U1), pc->du1, H >> 1, W >> 1)));
  // U2: b2 = U2(u2); u1 = a1 + up(b2)
  hipLaunchKernelGGL(k_up_t, dim3(ew_blocks(lvl(pc, 2, C1))), dim3(256), 0, st, pc->du1, pc->db2,
                     B, H >> 1, W >> 1, C1);
  WM_TRY((conv_dw<C2, C1>(ctx, pc, L_U2, pc->u2, pc->db2, H >> 2, W >> 2)));
  WM_TRY((conv<C1, C2, true>(ctx, pc->db2, wt + loff(L_U2), pc->du2, H >> 2, W >> 2)));
  // D3: a3 = D3(d3); u2 = a2 + up(a3)
  hipLaunchKernelGGL(k_up_t, dim3(ew_blocks(lvl(pc, 3, C2))), dim3(256), 0, st, pc->du2, pc->da3,
                     B, H >> 2, W >> 2, C2);
  WM_TRY((conv_dw<C2, C2>(ctx, pc, L_D3, pc->d3, pc->da3, H >> 3, W >> 3)));
  WM_TRY((conv<C2, C2, true>(ctx, pc->da3, wt + loff(L_D3), pc->dd3, H >> 3, W >> 3)));
  // D2: a2 = D2(d2); d3 = pool(a2)
  hipLaunchKernelGGL(k_add_unpool, dim3(ew_blocks(lvl(pc, 2, C2))), dim3(256), 0, st, pc->du2,
                     pc->dd3, pc->da2, B, H >> 2, W >> 2, C2);
  WM_TRY((conv_dw<C1, C2>(ctx, pc, L_D2, pc->d2, pc->da2, H >> 2, W >> 2)));
  WM_TRY((conv<C2, C1, true>(ctx, pc->da2, wt + loff(L_D2), pc->dd2, H >> 2, W >> 2)));
  // D1: a1 = D1(d1); d2 = pool(a1)
  hipLaunchKernelGGL(k_add_unpool, dim3(ew_blocks(lvl(pc, 1, C1))), dim3(256), 0, st, pc->du1,
                     pc->dd2, pc->da1, B, H >> 1, W >> 1, C1);
  WM_TRY((conv_dw<C0, C1>(ctx, pc, L_D1, pc->d1, pc->da1, H >> 1, W >> 1)));
  WM_TRY((conv<C1, C0, true>(ctx, pc->da1, wt + loff(L_D1), pc->dd1, H >> 1, W >> 1)));
  // lift: d1 = pool(a0), u0 = a0 + ...
  hipLaunchKernelGGL(k_lift_bwd, dim3(nb0), dim3(256), 0, st, r, stride, sc, pc->du0, pc->dd1,
                     pc->part, H, W, B, (int)t0.x, (int)t0.y, ntiles0);
  hipLaunchKernelGGL(k_reduce_rows, dim3(wm_cdiv(9 * C0, 32)), dim3(256), 0, st, pc->part, nb0,
                     9 * C0, pc->grad + loff(L_LIFT));
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int refresh_weights(wm_ctx *ctx, wm_pc *pc) {
  hipLaunchKernelGGL(k_weights_bf16, dim3(wm_cdiv(NPARAM, 256)), dim3(256), 0, ctx->stream,
                     pc->theta, pc->wb, pc->wt);
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, ctx->stream, pc->theta + loff(L_LIFT),
                     pc->theta + loff(L_PROJ), pc->fold);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

// --------------------------------------------------------------- launchers
int wm_pc_nparam() { return NPARAM; }

void wm_pc_free(wm_ctx *ctx) {
  wm_pc *pc = ctx->pc;
  if (!pc) return;
  void *ptrs[] = {pc->theta, pc->m,   pc->v,   pc->grad, pc->wb,  pc->wt,  pc->fold, pc->d1, pc->a1,
                  pc->d2,    pc->a2,  pc->d3,  pc->a3,   pc->u2,  pc->b2,  pc->u1,  pc->b1,
                  pc->y,     pc->du0, pc->du1, pc->du2,  pc->dd1, pc->dd2, pc->dd3, pc->db1,
                  pc->db2,   pc->da3, pc->da2, pc->da1,  pc->zt,  pc->part, pc->lossp,
                  pc->losses, pc->isc, pc->himg, pc->pr, pc->pe, pc->psc};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete pc;
  ctx->pc = nullptr;
}

// Allocates the net for the context's grid (nx, ny multiples of 8).
int wm_pc_alloc(wm_ctx *ctx) {
  if (ctx->pc) return WM_OK;
  const int nx = ctx->g.nx, ny = ctx->g.ny;
  if (nx % 8 || ny % 8) {
    wm_set_error(ctx, WM_ERR_ARG,
                 "learned preconditioner: nx=%d and ny=%d must be multiples of 8 (three 2x2 "
                 "coarsenings)", nx, ny);
    return WM_ERR_ARG;
  }
  wm_pc *pc = new wm_pc();
  ctx->pc = pc;
  pc->H = ny;
  pc->W = nx;
  const size_t n0 = (size_t)wm_pc::B * nx * ny;
  WM_TRY(dalloc(ctx, &pc->theta, NPARAM));
  WM_TRY(dalloc(ctx, &pc->m, NPARAM));
  WM_TRY(dalloc(ctx, &pc->v, NPARAM));
  WM_TRY(dalloc(ctx, &pc->grad, NPARAM));
  WM_TRY(dalloc(ctx, &pc->wb, NPARAM));
  WM_TRY(dalloc(ctx, &pc->wt, NPARAM));
  WM_TRY(dalloc(ctx, &pc->fold, NFOLD));
  WM_TRY(dalloc(ctx, &pc->d1, lvl(pc, 1, C0)));
  WM_TRY(dalloc(ctx, &pc->a1, lvl(pc, 1, C1)));
  WM_TRY(dalloc(ctx, &pc->d2, lvl(pc, 2, C1)));
  WM_TRY(dalloc(ctx, &pc->a2, lvl(pc, 2, C2)));
  WM_TRY(dalloc(ctx, &pc->d3, lvl(pc, 3, C2)));
  WM_TRY(dalloc(ctx, &pc->a3, lvl(pc, 3, C2)));
  WM_TRY(dalloc(ctx, &pc->u2, lvl(pc, 2, C2)));
  WM_TRY(dalloc(ctx, &pc->b2, lvl(pc, 2, C1)));
  WM_TRY(dalloc(ctx, &pc->u1, lvl(pc, 1, C1)));
  WM_TRY(dalloc(ctx, &pc->b1, lvl(pc, 1, C0)));
  WM_TRY(dalloc(ctx, &pc->isc, 6));
  WM_HIP(hipMemsetAsync(pc->m, 0, sizeof(float) * NPARAM, ctx->stream));
  WM_HIP(hipMemsetAsync(pc->v, 0, sizeof(float) * NPARAM, ctx->stream));
  (void)n0;
  return WM_OK;
}

static int alloc_training(wm_ctx *ctx, wm_pc *pc) {
  if (pc->y) return WM_OK;
  const size_t n0 = (size_t)wm_pc::B * pc->W * pc->H;
  WM_TRY(dalloc(ctx, &pc->y, n0));
  WM_TRY(dalloc(ctx, &pc->zt, n0));
  WM_TRY(dalloc(ctx, &pc->du0, lvl(pc, 0, C0)));
  WM_TRY(dalloc(ctx, &pc->du1, lvl(pc, 1, C1)));
  WM_TRY(dalloc(ctx, &pc->du2, lvl(pc, 2, C2)));
  WM_TRY(dalloc(ctx, &pc->dd1, lvl(pc, 1, C0)));
  WM_TRY(dalloc(ctx, &pc->dd2, lvl(pc, 2, C1)));
  WM_TRY(dalloc(ctx, &pc->dd3, lvl(pc, 3, C2)));
  WM_TRY(dalloc(ctx, &pc->db1, lvl(pc, 1, C0)));
  WM_TRY(dalloc(ctx, &pc->db2, lvl(pc, 2, C1)));
  WM_TRY(dalloc(ctx, &pc->da3, lvl(pc, 3, C2)));
  WM_TRY(dalloc(ctx, &pc->da2, lvl(pc, 2, C2)));
  WM_TRY(dalloc(ctx, &pc->da1, lvl(pc, 1, C1)));
  pc->nlossp = FINE_BLOCKS;
  pc->part_elems = std::max<size_t>((size_t)512 * MAXW, (size_t)FINE_BLOCKS * 9 * C0);
  WM_TRY(dalloc(ctx, &pc->part, pc->part_elems));
  WM_TRY(dalloc(ctx, &pc->lossp, pc->nlossp));
  WM_TRY(dalloc(ctx, &pc->himg, 2 * n0));
  return WM_OK;
}

int wm_pc_set_params_dev(wm_ctx *ctx, const float *theta_host, int n) {
  if (n != NPARAM) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: %d parameters given, %d expected", n,
                 NPARAM);
    return WM_ERR_ARG;
  }
  WM_TRY(wm_pc_alloc(ctx));
  wm_pc *pc = ctx->pc;
  WM_HIP(hipMemcpyAsync(pc->theta, theta_host, sizeof(float) * NPARAM, hipMemcpyHostToDevice,
                        ctx->stream));
  WM_HIP(hipMemsetAsync(pc->m, 0, sizeof(float) * NPARAM, ctx->stream));
  WM_HIP(hipMemsetAsync(pc->v, 0, sizeof(float) * NPARAM, ctx->stream));
  pc->adam_t = 0;
  WM_TRY(refresh_weights(ctx, pc));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  pc->trained = true;
  return WM_OK;
}

int wm_pc_get_params_dev(wm_ctx *ctx, float *theta_host, int n) {
  if (!ctx->pc || n != NPARAM) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: no parameters (or n != %d)", NPARAM);
    return WM_ERR_ARG;
  }
  WM_HIP(hipMemcpyAsync(theta_host, ctx->pc->theta, sizeof(float) * NPARAM, hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// Net applied to three host images r[3][ny][nx] -> z.
int wm_pc_apply_host(wm_ctx *ctx, const double *r, double *z) {
  wm_pc *pc = ctx->pc;
  if (!pc || !pc->trained) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: set or train the parameters first");
    return WM_ERR_ARG;
  }
  WM_TRY(alloc_training(ctx, pc));
  const size_t n = (size_t)pc->H * pc->W;
  WM_HIP(hipMemcpyAsync(pc->himg, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_img_scale, dim3(3), dim3(256), 0, ctx->stream, pc->himg, n, pc->isc);
  WM_TRY(forward(ctx, pc, pc->himg, n, pc->isc, pc->zt, n, nullptr));
  WM_HIP(hipMemcpyAsync(z, pc->zt, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// ---------------------------------------------- no-allocation application
// z = M(r) for callers that cannot allocate (code under graph capture, tight
// serving loops): the activations and the residual scales live in this file's
// device array `scratch`, sized for images of at most WM_PC_NOALLOC_CELLS
// pixels.  Each GPU holds its own copy of the array, at its own address,
// until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU race.  Nothing in it needs to start at zero: every
// kernel of the pass writes all of what the next one reads.
namespace {
constexpr size_t NOALLOC_CELLS = WM_PC_NOALLOC_CELLS;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// bytes of one activation at level l (pixels (ny >> l) (nx >> l) <= cells >> 2l)
constexpr size_t act_bytes(int l, int ch) {
  return align256((size_t)wm_pc::B * (NOALLOC_CELLS >> (2 * l)) * ch * sizeof(bf16));
}
constexpr size_t SC_BYTES = align256(6 * sizeof(double));
constexpr size_t SCRATCH_BYTES = SC_BYTES + act_bytes(1, C0) + act_bytes(1, C1) + act_bytes(1, C0) +
                                 act_bytes(2, C1) + act_bytes(2, C2) + act_bytes(2, C1) +
                                 act_bytes(3, C2) + act_bytes(3, C2);
}  // namespace

// In host code `scratch` names a host-side stand-in: kernels get its device
// address from hipGetSymbolAddress (scratch_of), never the name itself.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEV = 64;
void *scratch_ptr[MAX_DEV];            // the array's device address, per GPU (set up once)

int scratch_of(wm_ctx *ctx, unsigned char **out) {
  if (ctx->device < 0 || ctx->device >= MAX_DEV) {
    wm_set_error(ctx, WM_ERR_NODEV, "pc_apply_noalloc: device %d beyond %d", ctx->device, MAX_DEV);
    return WM_ERR_NODEV;
  }
  if (!scratch_ptr[ctx->device]) {
    int cur = 0;
    WM_HIP(hipGetDevice(&cur));
    WM_HIP(hipSetDevice(ctx->device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    WM_HIP(hipSetDevice(cur));
    WM_HIP(e);
    scratch_ptr[ctx->device] = p;
  }
  *out = (unsigned char *)scratch_ptr[ctx->device];
  return WM_OK;
}
}  // namespace

size_t wm_pc_noalloc_bytes() { return SCRATCH_BYTES; }

int wm_pc_apply_noalloc_dev(wm_ctx *ctx, const double *r, double *z, int nx, int ny) {
  wm_pc *pc = ctx->pc;
  if (!pc || !pc->trained) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: set or train the parameters first");
    return WM_ERR_ARG;
  }
  if (nx <= 0 || ny <= 0 || nx % 8 || ny % 8 || (size_t)nx * ny > NOALLOC_CELLS) {
    wm_set_error(ctx, WM_ERR_ARG,
                 "pc_apply_noalloc: nx=%d, ny=%d must be positive multiples of 8 with nx*ny <= %zu "
                 "(its scratch array's size)", nx, ny, NOALLOC_CELLS);
    return WM_ERR_ARG;
  }
  unsigned char *b = nullptr;
  WM_TRY(scratch_of(ctx, &b));
  double *sc = (double *)b;
  b += SC_BYTES;
  Acts X{};
  auto take = [&](int l, int ch) {
    bf16 *p = (bf16 *)b;
    b += act_bytes(l, ch);
    return p;
  };
  X.d1 = take(1, C0);
  X.a1 = take(1, C1);
  X.b1 = take(1, C0);
  X.d2 = take(2, C1);
  X.a2 = take(2, C2);
  X.b2 = take(2, C1);
  X.d3 = take(3, C2);
  X.a3 = take(3, C2);
  const size_t n = (size_t)nx * ny;
  hipLaunchKernelGGL(k_img_scale, dim3(3), dim3(256), 0, ctx->stream, r, n, sc);
  WM_HIP(hipGetLastError());
  WM_TRY(forward_on(ctx, pc, X, ny, nx, r, n, sc, z, n, nullptr));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// Loss and gradient for host pair images (r, e)[3][ny][nx].
int wm_pc_grad_host(wm_ctx *ctx, const double *r, const double *e, float *grad, double *loss) {
  wm_pc *pc = ctx->pc;
  if (!pc || !pc->trained) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: set or train the parameters first");
    return WM_ERR_ARG;
  }
  WM_TRY(alloc_training(ctx, pc));
  const size_t n = (size_t)pc->H * pc->W;
  WM_HIP(hipMemcpyAsync(pc->himg, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  WM_HIP(hipMemcpyAsync(pc->himg + 3 * n, e, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                        ctx->stream));
  hipLaunchKernelGGL(k_img_scale, dim3(3), dim3(256), 0, ctx->stream, pc->himg, n, pc->isc);
  double *ld = pc->lossp;   // reused: k_sum_loss reads lossp before writing here? no: separate slot
  if (pc->losses_cap < 1) {
    if (pc->losses) (void)hipFree(pc->losses);
    WM_TRY(dalloc(ctx, &pc->losses, 1024));
    pc->losses_cap = 1024;
  }
  ld = pc->losses;
  WM_TRY(backward(ctx, pc, pc->himg, pc->himg + 3 * n, n, pc->isc, ld));
  WM_HIP(hipMemcpyAsync(grad, pc->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipMemcpyAsync(loss, ld, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_pc_collect_cfg(wm_ctx *ctx, int nslot, int per_solve) {
  if (nslot < 0 || nslot > 64 || per_solve < 0) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: 0 <= slots <= 64, per_solve >= 0");
    return WM_ERR_ARG;
  }
  WM_TRY(wm_pc_alloc(ctx));
  wm_pc *pc = ctx->pc;
  const size_t n = (size_t)pc->H * pc->W;
  if (nslot > pc->nslot) {
    if (pc->pr) (void)hipFree(pc->pr);
    if (pc->pe) (void)hipFree(pc->pe);
    if (pc->psc) (void)hipFree(pc->psc);
    pc->pr = pc->pe = pc->psc = nullptr;
    WM_TRY(dalloc(ctx, &pc->pr, (size_t)nslot * 3 * n));
    WM_TRY(dalloc(ctx, &pc->pe, (size_t)nslot * 3 * n));
    WM_TRY(dalloc(ctx, &pc->psc, (size_t)nslot * 6));
    pc->filled = pc->next = 0;
  }
  pc->nslot = nslot;
  pc->per_solve = nslot ? per_solve : 0;
  if (pc->filled > nslot) pc->filled = nslot;
  if (pc->next >= std::max(nslot, 1)) pc->next = 0;
  return WM_OK;
}

bool wm_pc_collecting(const wm_ctx *ctx) { return ctx->pc && ctx->pc->per_solve > 0; }

// Called by the plain solver at the start of iteration k (all components
// still active).
int wm_pc_snapshot(wm_ctx *ctx, int k) {
  wm_pc *pc = ctx->pc;
  if (!pc || k >= pc->per_solve || pc->nsnap >= 64) return WM_OK;
  const int s = pc->next;
  pc->next = (pc->next + 1) % pc->nslot;
  if (pc->filled < pc->nslot) ++pc->filled;
  const size_t n = (size_t)pc->H * pc->W;
  hipLaunchKernelGGL(k_snap, dim3(std::min(wm_cdiv((long long)n, 256), 4096), 3), dim3(256), 0,
                     ctx->stream, ctx->cg_r, ctx->df, ctx->g.nx, ctx->g.ny, ctx->fg.gx, ctx->fg.ps,
                     ctx->cg_s, pc->pr + (size_t)s * 3 * n, pc->pe + (size_t)s * 3 * n,
                     pc->psc + (size_t)s * 6);
  WM_HIP(hipGetLastError());
  pc->snap_slot[pc->nsnap++] = s;
  return WM_OK;
}

// Called by the plain solver after convergence: corrections of the snapshots.
int wm_pc_snapshot_fin(wm_ctx *ctx) {
  wm_pc *pc = ctx->pc;
  if (!pc) return WM_OK;
  const size_t n = (size_t)pc->H * pc->W;
  for (int i = 0; i < pc->nsnap; ++i) {
    const int s = pc->snap_slot[i];
    hipLaunchKernelGGL(k_snap_fin, dim3(std::min(wm_cdiv((long long)n, 256), 4096), 3), dim3(256),
                       0, ctx->stream, ctx->df, ctx->g.nx, ctx->g.ny, ctx->fg.gx, ctx->fg.ps,
                       pc->pe + (size_t)s * 3 * n);
  }
  pc->nsnap = 0;
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int wm_pc_pairs(const wm_ctx *ctx) { return ctx->pc ? ctx->pc->filled : 0; }

// Default initial weights: the 1/f4 identity stencil through channel 0 of
// lift/proj plus small fan-in-scaled values (host counter-based generator).
static void init_theta(std::vector<float> &th, double f4, uint64_t seed) {
  th.assign(NPARAM, 0.0f);
  uint64_t s = seed ^ 0x243F6A8885A308D3ULL;
  for (int l = 0; l < NLAYER; ++l) {
    const int ci = LCIN[l], co = LCOUT[l];
    const double sd = 0.1 / sqrt(9.0 * ci);
    for (int k = 0; k < 9 * ci * co; ++k) {
      const double u1 = ((splitmix(s) >> 11) + 1) * 0x1.0p-53, u2 = (splitmix(s) >> 11) * 0x1.0p-53;
      th[loff(l) + k] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
  }
  th[loff(L_LIFT) + 4 * C0 + 0] = 1.0f;
  th[loff(L_PROJ) + 4 * C0 + 0] = (float)(1.0 / f4);
}

int wm_pc_init_params(wm_ctx *ctx, uint64_t seed) {
  WM_TRY(wm_pc_alloc(ctx));
  const double f5 = pow(ctx->g.delx / (ctx->g.c * ctx->g.delt * ctx->g.gfac), 2);
  std::vector<float> th;
  init_theta(th, 4.0 + f5, seed);
  return wm_pc_set_params_dev(ctx, th.data(), NPARAM);
}

// Adam training on the recorded pairs, cosine learning-rate decay from lr.
int wm_pc_train_dev(wm_ctx *ctx, int steps, double lr, uint64_t seed, double *loss_out) {
  wm_pc *pc = ctx->pc;
  if (!pc || pc->filled == 0) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: no (residual, correction) pairs recorded");
    return WM_ERR_ARG;
  }
  if (steps <= 0 || !(lr > 0)) {
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: steps > 0 and lr > 0 required");
    return WM_ERR_ARG;
  }
  if (!pc->trained) WM_TRY(wm_pc_init_params(ctx, seed));
  WM_TRY(alloc_training(ctx, pc));
  if (pc->losses_cap < steps) {
    if (pc->losses) (void)hipFree(pc->losses);
    pc->losses = nullptr;
    WM_TRY(dalloc(ctx, &pc->losses, steps));
    pc->losses_cap = steps;
  }
  const size_t n = (size_t)pc->H * pc->W;
  uint64_t rs = seed;
  const double b1 = 0.9, b2 = 0.999;
  for (int it = 0; it < steps; ++it) {
    const int s = (int)(splitmix(rs) % (uint64_t)pc->filled);
    WM_TRY(backward(ctx, pc, pc->pr + (size_t)s * 3 * n, pc->pe + (size_t)s * 3 * n, n,
                    pc->psc + (size_t)s * 6, pc->losses + it));
    ++pc->adam_t;
    const double lrt = lr * 0.5 * (1.0 + cos(M_PI * it / steps));
    const float c1 = (float)(1.0 / (1.0 - pow(b1, pc->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - pow(b2, pc->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(wm_cdiv(NPARAM, 256)), dim3(256), 0, ctx->stream, pc->theta,
                       pc->m, pc->v, pc->grad, NPARAM, (float)lrt, (float)b1, (float)b2, c1, c2,
                       1e-8f);
    WM_TRY(refresh_weights(ctx, pc));
  }
  std::vector<double> ls(steps);
  WM_HIP(hipMemcpyAsync(ls.data(), pc->losses, sizeof(double) * steps, hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  if (loss_out) {
    const int k = std::min(steps, 10);
    double a = 0.0;
    for (int i = steps - k; i < steps; ++i) a += ls[i];
    loss_out[0] = ls[0];
    loss_out[1] = a / k;
  }
  pc->trained = true;
  return WM_OK;
}

// Preconditioner application inside the solver: z[3][n] = M(r[3][n]) with
// device scales sc (s[3], 1/s[3]).
int wm_pc_apply_dev(wm_ctx *ctx, const double *r, const double *sc, double *z, const double *q,
                    double *dots, const FcgScal *S) {
  wm_pc *pc = ctx->pc;
  const size_t n = (size_t)pc->H * pc->W;
  return forward(ctx, pc, r, n, sc, z, n, nullptr, q, dots, S);
}

int wm_pc_dot_tiles(int nx, int ny) { return wm_cdiv(nx, T) * wm_cdiv(ny, T); }

bool wm_pc_ready(const wm_ctx *ctx) { return ctx->pc && ctx->pc->trained; }

void wm_pc_set_stats(wm_ctx *ctx, const int *fcg, const int *fb) {
  if (!ctx->pc) return;
  for (int l = 0; l < 3; ++l) {
    ctx->pc->last_fcg[l] = fcg[l];
    ctx->pc->last_fallback[l] = fb[l];
  }
}

void wm_pc_get_stats(const wm_ctx *ctx, int *fcg, int *fb) {
  for (int l = 0; l < 3; ++l) {
    fcg[l] = ctx->pc ? ctx->pc->last_fcg[l] : 0;
    fb[l] = ctx->pc ? ctx->pc->last_fallback[l] : 0;
  }
}
