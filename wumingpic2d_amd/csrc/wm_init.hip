// wm_init.hip -- synthetic counter-streaming pair plasma on the device.
//
// A counter-based generator keyed by (seed, particle index) makes every
// particle's draw independent of launch shape, so the host generator in
// wumingpic2d_amd/weibel.py produces the same state.  Positions are uniform
// within each row (x in [nxgs, nxge+1), y in [j, j+1)), every species uses
// the same positions (zero net charge, as proj/weibel/app.f90:410-411 pairs
// species), momenta are (0, 0, +/-udrift) + N(0, vth) per component with
// the drift sign alternating by particle index, IDs are -(global index + 1)
// as in proj/weibel/app.f90:462-468.  Particles are written unsorted to bank
// 1; the caller sorts them into bank 0.
#include "wm_internal.h"

namespace {

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ inline double uni(uint64_t seed, uint64_t ctr) {
  const uint64_t h = mix64(seed * 0xD1B54A32D192ED03ULL ^ mix64(ctr));
  return (double)(h >> 11) * 0x1.0p-53;
}
__device__ inline double gauss(uint64_t seed, uint64_t ctr) {
  const double u1 = 1.0 - uni(seed, 2 * ctr);       // (0, 1]
  const double u2 = uni(seed, 2 * ctr + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct IP {
  int nx, ny, nxgs, nys, nsp, cap, n;
  uint64_t seed;
  double vth, ud, xhi;   // xhi = largest double below nxge+1
};

// Stores one generated particle in the bank's form.
__device__ inline void put(const Rec64 &B, size_t p, double x, double y, double ux, double uy,
                           double uz, const IP &P, int j) {
  B.x[p] = x;
  B.y[p] = y;
  B.ux[p] = ux;
  B.uy[p] = uy;
  B.uz[p] = uz;
  (void)P;
  (void)j;
}
__device__ inline void put(const Rec32 &B, size_t p, double x, double y, double ux, double uy,
                           double uz, const IP &P, int j) {
  const int ix = (int)x;
  B.ci[p] = ix - P.nxgs;
  B.ox[p] = wm_below_one((float)(x - ix));
  B.oy[p] = wm_below_one((float)(y - j));
  B.ux[p] = (float)ux;
  B.uy[p] = (float)uy;
  B.uz[p] = (float)uz;
}

template <class R>
__global__ void k_init(IP P, R B, int64_t *id, int *np2) {
  const int seg = blockIdx.y;
  const int s = seg / P.ny, jl = seg % P.ny;
  const int j = P.nys + jl;
  const double yhi = __longlong_as_double(__double_as_longlong((double)(j + 1)) - 1);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P.n; k += gridDim.x * blockDim.x) {
    const uint64_t pidx = (uint64_t)jl * P.n + k;                 // species independent
    const uint64_t gidx = ((uint64_t)s * P.ny + jl) * P.n + k;
    double x = P.nxgs + P.nx * uni(P.seed, 4 * pidx);
    double y = j + uni(P.seed, 4 * pidx + 1);
    x = fmin(x, P.xhi);
    y = fmin(y, yhi);
    const uint64_t vs = P.seed ^ 0x5851F42D4C957F2DULL;
    const size_t p = (size_t)seg * P.cap + k;
    put(B, p, x, y, P.vth * gauss(vs, 3 * gidx), P.vth * gauss(vs, 3 * gidx + 1),
        ((k & 1) ? -P.ud : P.ud) + P.vth * gauss(vs, 3 * gidx + 2), P, j);
    id[p] = -(int64_t)(gidx + 1);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) np2[seg] = P.n;
}

}  // namespace

int wm_k_init_weibel(wm_ctx *ctx, uint64_t seed, int ppc, double vth, double udrift) {
  const wm_grid &g = ctx->g;
  IP P;
  P.nx = g.nx; P.ny = g.ny; P.nxgs = g.nxgs; P.nys = g.nys; P.nsp = g.nsp; P.cap = ctx->cap;
  const long long n = (long long)ppc * g.nx;
  if (n > ctx->cap) {
    wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np): %lld particles per row, capacity %d",
                 n, ctx->cap);
    return WM_ERR_NP_OVER;
  }
  P.n = (int)n;
  P.seed = seed;
  P.vth = vth;
  P.ud = udrift;
  P.xhi = nextafter((double)(g.nxge + 1), 0.0);
  const int nseg = g.nsp * g.ny;
  const int nb = std::max(1, std::min(64, wm_cdiv(n, 256)));
  if (ctx->form == WM_FORM_FP32)
    hipLaunchKernelGGL(k_init<Rec32>, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->b32[1],
                       ctx->id[ctx->idcur], ctx->np2);
  else
    hipLaunchKernelGGL(k_init<Rec64>, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->bank[1],
                       ctx->id[ctx->idcur], ctx->np2);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
