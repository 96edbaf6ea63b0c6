// wm_init.hip -- synthetic counter-streaming pair plasma on the device.
//
// A counter-based generator keyed by (seed, particle index) makes every
// particle's draw independent of launch shape, so the host generator in
// wumingpic2d_amd/weibel.py produces the same state.  Positions are uniform
// within each row (x in [nxgs, nxge+1), y in [j, j+1)), every species uses
// the same positions (zero net charge, as proj/weibel/app.f90:410-411 pairs
// species), momenta are (0, 0, +/-udrift) + N(0, vth) per component with
// the drift sign alternating by particle index, IDs are -(global index + 1)
// as in proj/weibel/app.f90:462-468.  The FP64 form's particles are written
// unsorted to bank 1 (the caller sorts them into bank 0); the FP32 form's go
// straight into the cell bins, in the order that sort would leave them.
#include "wm_bins.h"
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

// Particle k of row segment (s, jl): position (shared by the species) and
// momentum.
struct Gen {
  double x, y, ux, uy, uz;
  int64_t id;
};
__device__ inline double gen_x(const IP &P, int jl, int k) {
  const uint64_t pidx = (uint64_t)jl * P.n + k;                   // species independent
  return fmin(P.nxgs + P.nx * uni(P.seed, 4 * pidx), P.xhi);
}
__device__ inline Gen gen(const IP &P, int s, int jl, int k) {
  const int j = P.nys + jl;
  const double yhi = __longlong_as_double(__double_as_longlong((double)(j + 1)) - 1);
  const uint64_t pidx = (uint64_t)jl * P.n + k;
  const uint64_t gidx = ((uint64_t)s * P.ny + jl) * P.n + k;
  const uint64_t vs = P.seed ^ 0x5851F42D4C957F2DULL;
  Gen g;
  g.x = gen_x(P, jl, k);
  g.y = fmin(j + uni(P.seed, 4 * pidx + 1), yhi);
  g.ux = P.vth * gauss(vs, 3 * gidx);
  g.uy = P.vth * gauss(vs, 3 * gidx + 1);
  g.uz = ((k & 1) ? -P.ud : P.ud) + P.vth * gauss(vs, 3 * gidx + 2);
  g.id = -(int64_t)(gidx + 1);
  return g;
}

__global__ void k_init(IP P, Rec64 B, int64_t *id, int *np2) {
  const int seg = blockIdx.y;
  const int s = seg / P.ny, jl = seg % P.ny;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P.n; k += gridDim.x * blockDim.x) {
    const Gen v = gen(P, s, jl, k);
    const size_t p = (size_t)seg * P.cap + k;
    B.x[p] = v.x;
    B.y[p] = v.y;
    B.ux[p] = v.ux;
    B.uy[p] = v.uy;
    B.uz[p] = v.uz;
    id[p] = v.id;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) np2[seg] = P.n;
}

// FP32 form: the segment's particles straight into its cell bins, one
// wavefront per segment: the cell histogram, the layout, then the particles
// regenerated in row order and scattered stably (the order of sort__bucket).
__global__ __launch_bounds__(64) void k_init_bins(IP P, Bins32 A, int64_t *ids, int *bstart,
                                                  int *bin_cnt, int *derr) {
  extern __shared__ int run[];
  const int seg = blockIdx.x, lane = threadIdx.x;
  const int s = seg / P.ny, jl = seg % P.ny;
  const int j = P.nys + jl;
  int *bs = bstart + (size_t)seg * (P.nx + 1);
  for (int i = lane; i < P.nx; i += 64) run[i] = 0;
  __syncthreads();
  for (int k = lane; k < P.n; k += 64) atomicAdd(&run[(int)gen_x(P, jl, k) - P.nxgs], 1);
  __syncthreads();
  if (!wm_bins::wave_layout([&](int i) { return run[i]; }, P.nx, P.cap, bs)) {
    for (int i = lane; i <= P.nx; i += 64) bs[i] = 0;
    for (int i = lane; i < P.nx; i += 64) bin_cnt[(size_t)seg * P.nx + i] = 0;
    if (lane == 0) atomicOr(derr, DERR_NP_OVER);
    return;
  }
  for (int i = lane; i < P.nx; i += 64) {
    bin_cnt[(size_t)seg * P.nx + i] = run[i];
    run[i] = 0;
  }
  __syncthreads();                         // the layout's stores before its loads below
  for (int k0 = 0; k0 < P.n; k0 += 64) {
    const int k = k0 + lane;
    const bool act = k < P.n;
    const Gen v = gen(P, s, jl, act ? k : 0);
    const int ix = (int)v.x;
    const int key = act ? ix - P.nxgs : -1;
    const wm_bins::WaveMatch m = wm_bins::wave_match(key, act);
    if (act) {
      const size_t p = (size_t)seg * P.cap + bs[key] + run[key] + m.rank;
      A.q[p] = make_float4(wm_below_one((float)(v.x - ix)), wm_below_one((float)(v.y - j)),
                           (float)v.ux, (float)v.uy);
      A.uz[p] = (float)v.uz;
      if (ids) ids[p] = v.id;
    }
    __builtin_amdgcn_wave_barrier();
    if (act && m.leader) run[key] += m.count;
    __builtin_amdgcn_wave_barrier();
  }
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
  if (ctx->form == WM_FORM_FP32) {
    if (sizeof(int) * (size_t)g.nx > 65536) {
      wm_set_error(ctx, WM_ERR_ARG, "FP32 form: nx=%d exceeds the 16384 cells one LDS histogram "
                   "holds", g.nx);
      return WM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_init_bins, dim3(nseg), dim3(64), sizeof(int) * (size_t)g.nx, ctx->stream,
                       P, ctx->bins, ctx->id_bins, ctx->bstart, ctx->bin_cnt, ctx->derr);
  } else {
    hipLaunchKernelGGL(k_init, dim3(nb, nseg), dim3(256), 0, ctx->stream, P, ctx->bank[1],
                       ctx->id[ctx->idcur], ctx->np2);
  }
  WM_HIP(hipGetLastError());
  return WM_OK;
}
