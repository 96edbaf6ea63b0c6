// wm_api.hip -- the C ABI of libwmpic: contexts, module initialisers, the
// host-array entry points with the reference's call semantics, and the
// device-resident step.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "wm_internal.h"

size_t wm_deposit_part_elems(int nx, int ny);
size_t wm_deposit_part_elems32(int nx, int ny);
size_t wm_cg_partials(int nx, int ny);

void wm_set_error(wm_ctx *ctx, int code, const char *fmt, ...) {
  if (!ctx) return;
  ctx->status = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(ctx->msg, sizeof(ctx->msg), fmt, ap);
  va_end(ap);
}

namespace {

// --------------------------------------------------- layout conversion
// AoS staging rows (seg, k) of ndim doubles <-> SoA banks.
__global__ void k_aos_to_soa(int ndim, int w, int cap, const int *__restrict__ cnt,
                             const double *__restrict__ st, PBank B, int64_t *id) {
  const int seg = blockIdx.y;
  const int n = cnt[seg];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double *r = st + ((size_t)seg * w + k) * ndim;
    const size_t p = (size_t)seg * cap + k;
    B.x[p] = r[0];
    B.y[p] = r[1];
    B.ux[p] = r[2];
    B.uy[p] = r[3];
    B.uz[p] = r[4];
    if (ndim >= 6) id[p] = __double_as_longlong(r[5]);
  }
}
__global__ void k_soa_to_aos(int ndim, int w, int cap, const int *__restrict__ cnt,
                             double *__restrict__ st, PBank B, const int64_t *id) {
  const int seg = blockIdx.y;
  const int n = cnt[seg];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double *r = st + ((size_t)seg * w + k) * ndim;
    const size_t p = (size_t)seg * cap + k;
    r[0] = B.x[p];
    r[1] = B.y[p];
    r[2] = B.ux[p];
    r[3] = B.uy[p];
    r[4] = B.uz[p];
    if (ndim >= 6) r[5] = __longlong_as_double(id[p]);
  }
}
// Field arrays: component-fastest (ncomp, gx, gy) <-> component planes.
__global__ void k_aos_to_planes(int ncomp, size_t n, size_t ps, const double *__restrict__ a,
                                double *__restrict__ pl) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int c = 0; c < ncomp; ++c) pl[(size_t)c * ps + t] = a[t * ncomp + c];
}
__global__ void k_planes_to_aos(int ncomp, size_t n, size_t ps, double *__restrict__ a,
                                const double *__restrict__ pl) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int c = 0; c < ncomp; ++c) a[t * ncomp + c] = pl[(size_t)c * ps + t];
}
__global__ void k_fill_int(int *a, size_t n, int v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}
__global__ void k_fill_dbl(double *a, size_t n, double v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// ------------------------------------------------------------ helpers
int nseg_of(const wm_ctx *ctx) { return ctx->g.nsp * ctx->g.ny; }

int check_grid(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
               int nys, int nye, double delx, double delt, double c, const double *q,
               const double *r, bool have_phys) {
  if (ndim < 5 || ndim > 6 || np <= 0 || nsp <= 0 || nsp > WM_MAXSP || nxge < nxgs + 3 ||
      nyge < nygs || nye < nys) {
    wm_set_error(ctx, WM_ERR_ARG, "invalid grid: ndim=%d np=%d nsp=%d", ndim, np, nsp);
    return WM_ERR_ARG;
  }
  if (nys != nygs || nye != nyge) {
    wm_set_error(ctx, WM_ERR_ARG,
                 "one device context holds the whole y range (nys=nygs, nye=nyge): got %d:%d of %d:%d",
                 nys, nye, nygs, nyge);
    return WM_ERR_ARG;
  }
  if (have_phys && delx != 1.0) {
    // sort.f90:60,72 bins by int(x), i.e. assumes unit cells
    wm_set_error(ctx, WM_ERR_ARG, "delx must be 1 (sort__bucket bins particles by int(x))");
    return WM_ERR_ARG;
  }
  wm_grid g{};
  g.ndim = ndim; g.np = np; g.nsp = nsp;
  g.nxgs = nxgs; g.nxge = nxge; g.nygs = nygs; g.nyge = nyge; g.nys = nys; g.nye = nye;
  g.nx = nxge - nxgs + 1;
  g.ny = nye - nys + 1;
  if (ctx->have_grid) {
    const wm_grid &h = ctx->g;
    if (h.ndim != ndim || h.np != np || h.nsp != nsp || h.nxgs != nxgs || h.nxge != nxge ||
        h.nygs != nygs || h.nyge != nyge || h.nys != nys || h.nye != nye) {
      wm_set_error(ctx, WM_ERR_ARG, "module initialisers disagree on the grid");
      return WM_ERR_ARG;
    }
  } else {
    g.delx = 1.0;
    g.delt = 1.0;
    g.c = 1.0;
    g.gfac = 0.501;
    for (int s = 0; s < WM_MAXSP; ++s) g.r[s] = 1.0;
    ctx->g = g;
    ctx->have_grid = true;
  }
  if (have_phys) {
    ctx->g.delx = delx;
    ctx->g.delt = delt;
    ctx->g.c = c;
    for (int s = 0; s < nsp; ++s) {
      ctx->g.q[s] = q[s];
      ctx->g.r[s] = r[s];
    }
  }
  return WM_OK;
}

template <class T>
int dmalloc(wm_ctx *ctx, T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
  if (e != hipSuccess) {
    wm_set_error(ctx, WM_ERR_NOMEM, "device allocation of %zu bytes failed: %s", sizeof(T) * n,
                 hipGetErrorString(e));
    *p = nullptr;
    return WM_ERR_NOMEM;
  }
  return WM_OK;
}

int alloc_bank(wm_ctx *ctx, PBank &b, size_t n) {
  WM_TRY(dmalloc(ctx, &b.x, n));
  WM_TRY(dmalloc(ctx, &b.y, n));
  WM_TRY(dmalloc(ctx, &b.ux, n));
  WM_TRY(dmalloc(ctx, &b.uy, n));
  WM_TRY(dmalloc(ctx, &b.uz, n));
  return WM_OK;
}
void free_bank(PBank &b) {
  (void)hipFree(b.x);
  (void)hipFree(b.y);
  (void)hipFree(b.ux);
  (void)hipFree(b.uy);
  (void)hipFree(b.uz);
  b = PBank{};
}

int fill_zero(wm_ctx *ctx, double *a, size_t n) {
  hipLaunchKernelGGL(k_fill_dbl, dim3(wm_cdiv((long long)n, 256)), dim3(256), 0, ctx->stream, a,
                     n, 0.0);
  WM_HIP(hipGetLastError());
  return WM_OK;
}

int ensure_stage(wm_ctx *ctx, size_t elems) {
  if (ctx->stage_elems >= elems) return WM_OK;
  if (ctx->stage) (void)hipFree(ctx->stage);
  ctx->stage = nullptr;
  ctx->stage_elems = 0;
  WM_TRY(dmalloc(ctx, &ctx->stage, elems));
  ctx->stage_elems = elems;
  return WM_OK;
}

// Upload the active records (cnt[seg] per segment) of a reference-layout
// particle array into bank `b` and the current ID array.
int upload_particles(wm_ctx *ctx, const double *up, const int *cnt_host, int b) {
  const wm_grid &g = ctx->g;
  const int nseg = nseg_of(ctx);
  int w = 0;
  for (int s = 0; s < nseg; ++s) {
    if (cnt_host[s] < 0 || cnt_host[s] > g.np) {
      wm_set_error(ctx, WM_ERR_RANGE, "particle count %d of row segment %d outside 0..np=%d",
                   cnt_host[s], s, g.np);
      return WM_ERR_RANGE;
    }
    if (cnt_host[s] > ctx->cap) {
      wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np): %d particles, device capacity %d",
                   cnt_host[s], ctx->cap);
      return WM_ERR_NP_OVER;
    }
    w = std::max(w, cnt_host[s]);
  }
  w = std::max(w, 1);
  const size_t elems = (size_t)nseg * w * g.ndim;
  WM_TRY(ensure_stage(ctx, elems));
  std::vector<double> h(elems);
  for (int s = 0; s < nseg; ++s)
    memcpy(&h[(size_t)s * w * g.ndim], up + (size_t)s * g.np * g.ndim,
           sizeof(double) * (size_t)cnt_host[s] * g.ndim);
  int *dcnt = ctx->cnt_scratch;
  WM_HIP(hipMemcpyAsync(ctx->stage, h.data(), sizeof(double) * elems, hipMemcpyHostToDevice,
                        ctx->stream));
  WM_HIP(hipMemcpyAsync(dcnt, cnt_host, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_aos_to_soa, dim3(std::min(64, wm_cdiv(w, 256)), nseg), dim3(256), 0,
                     ctx->stream, g.ndim, w, ctx->cap, dcnt, ctx->stage, ctx->bank[b],
                     ctx->id[ctx->idcur]);
  WM_HIP(hipGetLastError());
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// Write the active records of bank `b` (cnt_host per segment) into a
// reference-layout host array; other slots are left untouched.
int download_particles(wm_ctx *ctx, double *up, const int *cnt_host, int b, int ncomp = 0) {
  const wm_grid &g = ctx->g;
  const int nseg = nseg_of(ctx);
  int w = 1;
  for (int s = 0; s < nseg; ++s) w = std::max(w, cnt_host[s]);
  if (w > g.np) {
    wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np): %d > %d", w, g.np);
    return WM_ERR_NP_OVER;
  }
  const size_t elems = (size_t)nseg * w * g.ndim;
  WM_TRY(ensure_stage(ctx, elems));
  int *dcnt = ctx->cnt_scratch;
  WM_HIP(hipMemcpyAsync(dcnt, cnt_host, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_soa_to_aos, dim3(std::min(64, wm_cdiv(w, 256)), nseg), dim3(256), 0,
                     ctx->stream, g.ndim, w, ctx->cap, dcnt, ctx->stage, ctx->bank[b],
                     ctx->id[ctx->idcur]);
  WM_HIP(hipGetLastError());
  std::vector<double> h(elems);
  WM_HIP(hipMemcpyAsync(h.data(), ctx->stage, sizeof(double) * elems, hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  if (ncomp > 0 && ncomp < g.ndim) {      // the first ncomp components of each record only
    for (int s = 0; s < nseg; ++s)
      for (int k = 0; k < cnt_host[s]; ++k)
        memcpy(up + ((size_t)s * g.np + k) * g.ndim, &h[((size_t)s * w + k) * g.ndim],
               sizeof(double) * ncomp);
    return WM_OK;
  }
  for (int s = 0; s < nseg; ++s)
    memcpy(up + (size_t)s * g.np * g.ndim, &h[(size_t)s * w * g.ndim],
           sizeof(double) * (size_t)cnt_host[s] * g.ndim);
  return WM_OK;
}

// FP32 form: the reference's rows <-> cell bins, through the staging buffer
// in batches of segments of at most WM_STAGE_ELEMS doubles (at least one
// segment), so that no dense copy of the state is kept on the device.
#define WM_STAGE_ELEMS ((size_t)1 << 26)
}  // namespace

int wm_batch_end(const wm_ctx *ctx, const int *cnt, int s0, int *wout) {
  int &w = *wout;
  const int nseg = nseg_of(ctx);
  w = std::max(1, cnt[s0]);
  int s1 = s0 + 1;
  while (s1 < nseg) {
    const int w1 = std::max(w, cnt[s1]);
    if ((size_t)(s1 + 1 - s0) * w1 * ctx->g.ndim > WM_STAGE_ELEMS) break;
    w = w1;
    ++s1;
  }
  return s1;
}

namespace {
int upload_bins(wm_ctx *ctx, const double *up, const int *cnt_host) {
  const wm_grid &g = ctx->g;
  const int nseg = nseg_of(ctx);
  for (int s = 0; s < nseg; ++s)
    if (cnt_host[s] < 0 || cnt_host[s] > g.np) {
      wm_set_error(ctx, WM_ERR_RANGE, "particle count %d of row segment %d outside 0..np=%d",
                   cnt_host[s], s, g.np);
      return WM_ERR_RANGE;
    }
  for (int s0 = 0; s0 < nseg;) {
    int w = 1;
    const int s1 = wm_batch_end(ctx, cnt_host, s0, &w);
    const size_t elems = (size_t)(s1 - s0) * w * g.ndim;
    WM_TRY(ensure_stage(ctx, elems));
    std::vector<double> h(elems);
    for (int s = s0; s < s1; ++s)
      memcpy(&h[(size_t)(s - s0) * w * g.ndim], up + (size_t)s * g.np * g.ndim,
             sizeof(double) * (size_t)cnt_host[s] * g.ndim);
    WM_HIP(hipMemcpyAsync(ctx->stage, h.data(), sizeof(double) * elems, hipMemcpyHostToDevice,
                          ctx->stream));
    WM_HIP(hipMemcpyAsync(ctx->cnt_scratch, cnt_host + s0, sizeof(int) * (s1 - s0),
                          hipMemcpyHostToDevice, ctx->stream));
    WM_TRY(wm_k_stage_to_bins(ctx, s0, s1 - s0, w));
    WM_HIP(hipStreamSynchronize(ctx->stream));
    s0 = s1;
  }
  return WM_OK;
}

// The bins' particles (row counts cnt_host, as k_bin_rows left them) into
// the reference-layout host array up, sorted by cell within each row.
int download_bins(wm_ctx *ctx, double *up, const int *cnt_host) {
  const wm_grid &g = ctx->g;
  const int nseg = nseg_of(ctx);
  for (int s = 0; s < nseg; ++s)
    if (cnt_host[s] > g.np) {
      wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np): %d > %d", cnt_host[s], g.np);
      return WM_ERR_NP_OVER;
    }
  for (int s0 = 0; s0 < nseg;) {
    int w = 1;
    const int s1 = wm_batch_end(ctx, cnt_host, s0, &w);
    const size_t elems = (size_t)(s1 - s0) * w * g.ndim;
    WM_TRY(ensure_stage(ctx, elems));
    WM_TRY(wm_k_bins_to_stage(ctx, s0, s1 - s0, w));
    std::vector<double> h(elems);
    WM_HIP(hipMemcpyAsync(h.data(), ctx->stage, sizeof(double) * elems, hipMemcpyDeviceToHost,
                          ctx->stream));
    WM_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = s0; s < s1; ++s)
      memcpy(up + (size_t)s * g.np * g.ndim, &h[(size_t)(s - s0) * w * g.ndim],
             sizeof(double) * (size_t)cnt_host[s] * g.ndim);
    s0 = s1;
  }
  return WM_OK;
}

// FP32 form: a fresh state's bookkeeping (no pending movers, statistics
// cleared).
int fresh_bins(wm_ctx *ctx) {
  hipLaunchKernelGGL(k_fill_int, dim3(wm_cdiv(ctx->g.ny, 256)), dim3(256), 0, ctx->stream,
                     ctx->rowpend, (size_t)ctx->g.ny, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, ctx->stream, ctx->fstat, (size_t)WM_NSTAT, 0);
  WM_HIP(hipGetLastError());
  ctx->mv_pending = false;
  return WM_OK;
}

// Field array (ncomp, gx, gy) host <-> device planes.
int upload_field(wm_ctx *ctx, const double *a, int ncomp, double *planes) {
  const size_t n = (size_t)ctx->fg.gx * ctx->fg.gy;
  WM_TRY(ensure_stage(ctx, n * ncomp));
  WM_HIP(hipMemcpyAsync(ctx->stage, a, sizeof(double) * n * ncomp, hipMemcpyHostToDevice,
                        ctx->stream));
  hipLaunchKernelGGL(k_aos_to_planes, dim3(wm_cdiv((long long)n, 256)), dim3(256), 0, ctx->stream,
                     ncomp, n, ctx->fg.ps, ctx->stage, planes);
  WM_HIP(hipGetLastError());
  return WM_OK;
}
int download_field(wm_ctx *ctx, double *a, int ncomp, const double *planes) {
  const size_t n = (size_t)ctx->fg.gx * ctx->fg.gy;
  WM_TRY(ensure_stage(ctx, n * ncomp));
  hipLaunchKernelGGL(k_planes_to_aos, dim3(wm_cdiv((long long)n, 256)), dim3(256), 0, ctx->stream,
                     ncomp, n, ctx->fg.ps, ctx->stage, planes);
  WM_HIP(hipGetLastError());
  WM_HIP(hipMemcpyAsync(a, ctx->stage, sizeof(double) * n * ncomp, hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// cumcnt as the device reads it: the wall families' entries outside
// nxs .. nxe+1 are whatever the last sort left there (sort.f90:66-69 writes
// only those), so the cells outside the active range are made empty.
std::vector<int> active_cumcnt(const wm_ctx *ctx, const int *cumcnt) {
  const int nseg = nseg_of(ctx), w = ctx->g.nx + 1;
  std::vector<int> c(cumcnt, cumcnt + (size_t)nseg * w);
  if (ctx->bcf == WM_BC_PERIODIC || ctx->xe == 0) return c;
  const int a = ctx->xs - ctx->g.nxgs, b = ctx->xe + 1 - ctx->g.nxgs;
  for (int s = 0; s < nseg; ++s) {
    int *r = c.data() + (size_t)s * w;
    for (int i = 0; i < a; ++i) r[i] = 0;
    for (int i = b + 1; i < w; ++i) r[i] = r[b];
  }
  return c;
}

int upload_cumcnt(wm_ctx *ctx, const int *cumcnt_in) {
  const size_t n = (size_t)nseg_of(ctx) * (ctx->g.nx + 1);
  const std::vector<int> act = active_cumcnt(ctx, cumcnt_in);
  const int *cumcnt = act.data();
  // cumcnt gives cell i the particles cumcnt(i)+1 .. cumcnt(i+1)
  // (particle.f90:94, field.f90:218): inside the row capacity, and the
  // non-empty ranges in order without overlap.  A range may be empty with
  // cumcnt(i+1) < cumcnt(i): the shock's inject() appends particles to cell
  // nxe-1 by raising cumcnt(nxe) alone (proj/shock/app.f90:832-835).
  const int lim = std::min(ctx->g.np, ctx->cap);
  for (int s = 0; s < nseg_of(ctx); ++s) {
    const int *c = cumcnt + (size_t)s * (ctx->g.nx + 1);
    bool ok = true;
    int end = 0;
    for (int i = 0; ok && i <= ctx->g.nx; ++i) ok = c[i] >= 0 && c[i] <= lim;
    for (int i = 0; ok && i < ctx->g.nx; ++i)
      if (c[i + 1] > c[i]) {
        ok = c[i] >= end;
        end = c[i + 1];
      }
    if (!ok) {
      wm_set_error(ctx, WM_ERR_RANGE, "cumcnt of row segment %d is not a prefix table within capacity", s);
      return WM_ERR_RANGE;
    }
  }
  WM_HIP(hipMemcpyAsync(ctx->cumcnt, cumcnt, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  // the copy reads act before this scope ends
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// The particles a cumcnt table hands out: the largest entry of the active
// range (the end of its last non-empty cell range).
std::vector<int> counts_from_cumcnt(const wm_ctx *ctx, const int *cumcnt) {
  const int nseg = nseg_of(ctx);
  const bool all = ctx->bcf == WM_BC_PERIODIC || ctx->xe == 0;
  const int a = all ? 0 : ctx->xs - ctx->g.nxgs, b = all ? ctx->g.nx : ctx->xe + 1 - ctx->g.nxgs;
  std::vector<int> c(nseg);
  for (int s = 0; s < nseg; ++s) {
    const int *r = cumcnt + (size_t)s * (ctx->g.nx + 1);
    c[s] = *std::max_element(r + a, r + b + 1);
  }
  return c;
}

int require(wm_ctx *ctx, bool flag, const char *what) {
  if (!flag) {
    wm_set_error(ctx, WM_ERR_NOT_INIT, "Initialize first by calling %s()", what);
    return WM_ERR_NOT_INIT;
  }
  return WM_OK;
}

int require_fp64(wm_ctx *ctx, const char *what) {
  if (ctx->form != WM_FORM_FP64) {
    wm_set_error(ctx, WM_ERR_ARG, "%s works on the FP64 particle form (reference arithmetic)", what);
    return WM_ERR_ARG;
  }
  return WM_OK;
}

// The host-array entry points stage their arguments through the particle
// banks, fields, np2 and cumcnt that a device-resident state lives in.
int require_no_resident(wm_ctx *ctx, const char *what) {
  if (ctx->resident) {
    wm_set_error(ctx, WM_ERR_ARG,
                 "%s (host arrays) would overwrite this context's device-resident state; call it "
                 "on a context without one",
                 what);
    return WM_ERR_ARG;
  }
  return WM_OK;
}

// The active x range of a call: the whole grid on one periodic rank; the
// wall families' [nxs, nxe] inside it (the shock's box grows,
// proj/shock/app.f90:613-623), kept for the device kernels.
int check_bounds(wm_ctx *ctx, int nxs, int nxe) {
  if (ctx->bcf == WM_BC_PERIODIC) {
    if (nxs != ctx->g.nxgs || nxe != ctx->g.nxge) {
      wm_set_error(ctx, WM_ERR_ARG, "nxs:nxe = %d:%d must equal nxgs:nxge = %d:%d on one rank",
                   nxs, nxe, ctx->g.nxgs, ctx->g.nxge);
      return WM_ERR_ARG;
    }
    return WM_OK;
  }
  if (nxs < ctx->g.nxgs || nxe > ctx->g.nxge || nxe < nxs + 2) {
    wm_set_error(ctx, WM_ERR_ARG, "nxs:nxe = %d:%d must lie in nxgs:nxge = %d:%d, at least 3 cells",
                 nxs, nxe, ctx->g.nxgs, ctx->g.nxge);
    return WM_ERR_ARG;
  }
  ctx->xs = nxs;
  ctx->xe = nxe;
  return WM_OK;
}

}  // namespace

// ------------------------------------------ the state row by row (restart I/O)
int wm_state_np2(wm_ctx *ctx, int *np2_host) {
  if (ctx->form == WM_FORM_FP32) {
    WM_TRY(wm_settle_movers(ctx));
    WM_TRY(wm_k_bin_rows(ctx));
  }
  WM_HIP(hipMemcpyAsync(np2_host, ctx->np2, sizeof(int) * nseg_of(ctx), hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return wm_check_derr(ctx, "state rows");
}

int wm_rows_out(wm_ctx *ctx, int s0, int s1, int w, double *h) {
  const wm_grid &g = ctx->g;
  const size_t elems = (size_t)(s1 - s0) * w * g.ndim;
  WM_TRY(ensure_stage(ctx, elems));
  if (ctx->form == WM_FORM_FP32) {
    WM_TRY(wm_k_bins_to_stage(ctx, s0, s1 - s0, w));
  } else {
    const size_t off = (size_t)s0 * ctx->cap;
    const PBank &A = ctx->bank[0];
    const PBank B{A.x + off, A.y + off, A.ux + off, A.uy + off, A.uz + off};
    hipLaunchKernelGGL(k_soa_to_aos, dim3(std::min(64, wm_cdiv(w, 256)), s1 - s0), dim3(256), 0,
                       ctx->stream, g.ndim, w, ctx->cap, ctx->np2 + s0, ctx->stage, B,
                       ctx->id[ctx->idcur] + off);
    WM_HIP(hipGetLastError());
  }
  WM_HIP(hipMemcpyAsync(h, ctx->stage, sizeof(double) * elems, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_rows_in(wm_ctx *ctx, int s0, int s1, int w, const int *cnt_host, const double *h) {
  const wm_grid &g = ctx->g;
  const size_t elems = (size_t)(s1 - s0) * w * g.ndim;
  WM_TRY(ensure_stage(ctx, elems));
  WM_HIP(hipMemcpyAsync(ctx->stage, h, sizeof(double) * elems, hipMemcpyHostToDevice, ctx->stream));
  WM_HIP(hipMemcpyAsync(ctx->cnt_scratch, cnt_host + s0, sizeof(int) * (s1 - s0),
                        hipMemcpyHostToDevice, ctx->stream));
  if (ctx->form == WM_FORM_FP32) {
    WM_TRY(wm_k_stage_to_bins(ctx, s0, s1 - s0, w));
  } else {
    const size_t off = (size_t)s0 * ctx->cap;
    const PBank &A = ctx->bank[1];
    const PBank B{A.x + off, A.y + off, A.ux + off, A.uy + off, A.uz + off};
    hipLaunchKernelGGL(k_aos_to_soa, dim3(std::min(64, wm_cdiv(w, 256)), s1 - s0), dim3(256), 0,
                       ctx->stream, g.ndim, w, ctx->cap, ctx->cnt_scratch, ctx->stage, B,
                       ctx->id[ctx->idcur] + off);
    WM_HIP(hipGetLastError());
  }
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_rows_begin(wm_ctx *ctx) {
  WM_TRY(wm_alloc_device(ctx));
  if (ctx->form == WM_FORM_FP32) WM_TRY(fresh_bins(ctx));
  // a restarted run begins as a fresh process does: the CG warm start df is
  // zero (field.f90:105-119)
  WM_TRY(fill_zero(ctx, ctx->df, 6 * ctx->fg.ps));
  ctx->x0_fresh = false;
  return WM_OK;
}

int wm_rows_done(wm_ctx *ctx, const int *np2_host) {
  if (ctx->form == WM_FORM_FP64) {
    // the rows as read (bank 1), sorted by cell into bank 0 with cumcnt: the
    // reference's io__input then sort__bucket (proj/weibel/app.f90:351-352)
    WM_HIP(hipMemcpyAsync(ctx->np2, np2_host, sizeof(int) * nseg_of(ctx), hipMemcpyHostToDevice,
                          ctx->stream));
    WM_TRY(wm_k_sort(ctx, 1, 0));
  }
  WM_HIP(hipStreamSynchronize(ctx->stream));
  WM_TRY(wm_check_derr(ctx, "state rows"));
  ctx->resident = true;
  return WM_OK;
}

int wm_field_out(wm_ctx *ctx, double *uf) { return download_field(ctx, uf, 6, ctx->uf); }
int wm_field_in(wm_ctx *ctx, const double *uf) {
  WM_TRY(upload_field(ctx, uf, 6, ctx->uf));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// ------------------------------------------------------------- allocation
int wm_alloc_device(wm_ctx *ctx) {
  if (ctx->alloc_done) return WM_OK;
  if (!ctx->have_grid) {
    wm_set_error(ctx, WM_ERR_NOT_INIT, "Initialize first by calling particle__init()");
    return WM_ERR_NOT_INIT;
  }
  const wm_grid &g = ctx->g;
  if (ctx->cap <= 0) ctx->cap = g.np;
  ctx->cap = (ctx->cap + 15) / 16 * 16;   // segments start on 128-byte lines
  // the FP64 push addresses a segment's doubles by 32-bit byte offsets
  if (ctx->cap >= (1 << 29)) {
    wm_set_error(ctx, WM_ERR_ARG, "%d slots per row segment exceed 2^29", ctx->cap);
    return WM_ERR_ARG;
  }
  if (ctx->form == WM_FORM_FP64 && ctx->mcap <= 0) ctx->mcap = std::max(1024, ctx->cap / 8);
  const int nseg = g.nsp * g.ny;
  const size_t slots = (size_t)nseg * ctx->cap;
  if (ctx->form == WM_FORM_FP32) {
    // cell bins inside the row segments (wm_internal.h); the row pools hold a
    // quarter of a row's slots by default
    const size_t nbin = (size_t)nseg * g.nx;
    if (nbin > 0xFFFFFFFFu) {
      wm_set_error(ctx, WM_ERR_ARG, "FP32 form: %zu cell bins exceed 2^32", nbin);
      return WM_ERR_ARG;
    }
    const long long mrow = ctx->mcap > 0 ? ctx->mcap : std::max(1024LL, (long long)g.nsp * ctx->cap / 4);
    if (mrow >= (1LL << 24)) {
      wm_set_error(ctx, WM_ERR_ARG, "FP32 form: %lld mover records per row exceed 2^24", mrow);
      return WM_ERR_ARG;
    }
    ctx->mrow = (int)((mrow + 15) / 16 * 16);
    WM_TRY(dmalloc(ctx, &ctx->bins.q, slots));
    WM_TRY(dmalloc(ctx, &ctx->bins.uz, slots));
    if (g.ndim >= 6) WM_TRY(dmalloc(ctx, &ctx->id_bins, slots));
    WM_TRY(dmalloc(ctx, &ctx->bstart, (size_t)nseg * (g.nx + 1)));
    WM_TRY(dmalloc(ctx, &ctx->bin_cnt, nbin));
    WM_TRY(dmalloc(ctx, &ctx->pcnt, nbin));
    for (int b = 0; b < 2; ++b) {
      WM_TRY(dmalloc(ctx, &ctx->bin_sum[b], nbin));
      WM_TRY(dmalloc(ctx, &ctx->mv[b].rec, (size_t)g.ny * ctx->mrow));
      WM_TRY(dmalloc(ctx, &ctx->mv[b].perm, (size_t)g.ny * ctx->mrow));
      WM_TRY(dmalloc(ctx, &ctx->mfill[b], g.ny));
      WM_HIP(hipMemsetAsync(ctx->mfill[b], 0, sizeof(int) * g.ny, ctx->stream));
    }
    // the spill of the push's mover lists: a sixteenth of a row's slots
    // (block indices of the lists are 16-bit: at most 65535 - 60 spill blocks per row)
    ctx->srow = (int)std::min(16LL * 65000, (std::max(4096LL, (long long)g.nsp * ctx->cap / 16) + 15) / 16 * 16);
    const size_t sn = (size_t)g.ny * ctx->srow;
    WM_TRY(dmalloc(ctx, &ctx->spill_l, sn));
    WM_TRY(dmalloc(ctx, &ctx->spill_p, sn));
    WM_TRY(dmalloc(ctx, &ctx->spill_q, sn));
    WM_TRY(dmalloc(ctx, &ctx->sfill, g.ny));
    WM_HIP(hipMemsetAsync(ctx->sfill, 0, sizeof(int) * g.ny, ctx->stream));
    WM_TRY(dmalloc(ctx, &ctx->rowpend, g.ny));
    WM_TRY(dmalloc(ctx, &ctx->fstat, WM_NSTAT));
    // the re-layout's lists and workspace: relay_nwb block workgroups (2048:
    // eight per CU; a step at 8192 x 8192 cells, 32 per cell and species,
    // re-lays about 270k blocks) of relay_wsz
    // slots (four times a block's mean span), or relay_nws (256: one per CU)
    // segment workgroups of a segment's slots
    const size_t nblk = (size_t)nseg * wm_cdiv(g.nx, WM_RB);
    WM_TRY(dmalloc(ctx, &ctx->relay_ctl, 4));
    WM_TRY(dmalloc(ctx, &ctx->relay_blocks, nblk));
    WM_TRY(dmalloc(ctx, &ctx->blkflag, nblk));
    WM_TRY(dmalloc(ctx, &ctx->relay_segs, nseg));
    WM_TRY(dmalloc(ctx, &ctx->segflag, nseg));
    ctx->relay_wsz = std::min(ctx->cap, std::max(1024, 4 * WM_RB * wm_cdiv(ctx->cap, g.nx) + 63) / 64 * 64);
    ctx->relay_nwb = (int)std::min<size_t>(2048, nblk);
    ctx->relay_nws = std::min(256, nseg);
    const size_t wslots = std::max((size_t)ctx->relay_nwb * ctx->relay_wsz,
                                   (size_t)ctx->relay_nws * ctx->cap);
    WM_TRY(dmalloc(ctx, &ctx->rws.q, wslots));
    WM_TRY(dmalloc(ctx, &ctx->rws.uz, wslots));
    if (g.ndim >= 6) WM_TRY(dmalloc(ctx, &ctx->rws_id, wslots));
    WM_TRY(dmalloc(ctx, &ctx->rws_bs, (size_t)ctx->relay_nws * 2 * (g.nx + 1)));
    WM_HIP(hipMemsetAsync(ctx->relay_ctl, 0, sizeof(int) * 4, ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->blkflag, 0, sizeof(int) * nblk, ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->segflag, 0, sizeof(int) * nseg, ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->bin_cnt, 0, sizeof(int) * nbin, ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->bstart, 0, sizeof(int) * (size_t)nseg * (g.nx + 1), ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->rowpend, 0, sizeof(int) * g.ny, ctx->stream));
    WM_HIP(hipMemsetAsync(ctx->fstat, 0, sizeof(int) * WM_NSTAT, ctx->stream));
    ctx->mv_pending = false;
  } else {
    WM_TRY(alloc_bank(ctx, ctx->bank[0], slots));
    WM_TRY(alloc_bank(ctx, ctx->bank[1], slots));
    WM_TRY(dmalloc(ctx, &ctx->id[0], slots));
    WM_TRY(dmalloc(ctx, &ctx->id[1], slots));
  }
  WM_TRY(dmalloc(ctx, &ctx->np2, nseg));
  WM_TRY(dmalloc(ctx, &ctx->cumcnt, (size_t)nseg * (g.nx + 1)));
  if (ctx->form == WM_FORM_FP64) {
    // particle_y workspace
    const size_t ms = (size_t)nseg * ctx->mcap;
    WM_TRY(dmalloc(ctx, &ctx->hole, 2 * ms));
    WM_TRY(dmalloc(ctx, &ctx->nhole, nseg));
    WM_TRY(dmalloc(ctx, &ctx->nup, nseg));
    WM_TRY(dmalloc(ctx, &ctx->ndn, nseg));
    WM_TRY(alloc_bank(ctx, ctx->mv_up, ms));
    WM_TRY(alloc_bank(ctx, ctx->mv_dn, ms));
    WM_TRY(dmalloc(ctx, &ctx->mvid_up, ms));
    WM_TRY(dmalloc(ctx, &ctx->mvid_dn, ms));
    ctx->ytiles = wm_cdiv(ctx->cap, 64);
    WM_TRY(dmalloc(ctx, &ctx->ytmask, (size_t)nseg * ctx->ytiles));
    WM_TRY(dmalloc(ctx, &ctx->ytpre, (size_t)nseg * ctx->ytiles));
    WM_TRY(dmalloc(ctx, &ctx->yhist, (size_t)nseg * g.nx));
    WM_HIP(hipMemsetAsync(ctx->yhist, 0, sizeof(int) * (size_t)nseg * g.nx, ctx->stream));
  }

  FieldGeom &fg = ctx->fg;
  fg.nx = g.nx;
  fg.ny = g.ny;
  fg.gx = g.nx + 4;
  fg.gy = g.ny + 4;
  fg.ps = ((size_t)fg.gx * fg.gy + 63) / 64 * 64;
  WM_TRY(dmalloc(ctx, &ctx->uf, 6 * fg.ps));
  WM_TRY(dmalloc(ctx, &ctx->df, 6 * fg.ps));
  WM_TRY(dmalloc(ctx, &ctx->uj, 3 * fg.ps));
  if (ctx->form == WM_FORM_FP32)
    WM_TRY(dmalloc(ctx, &ctx->tmp32, 6 * (size_t)(g.nx + 2) * (g.ny + 2)));
  else
    WM_TRY(dmalloc(ctx, &ctx->tmp, 6 * (size_t)(g.nx + 2) * (g.ny + 2)));
  const size_t n = (size_t)g.nx * g.ny;
  WM_TRY(dmalloc(ctx, &ctx->gkl, 3 * n));
  WM_TRY(dmalloc(ctx, &ctx->cg_r, 3 * n));
  WM_TRY(dmalloc(ctx, &ctx->cg_p[0], 3 * n));
  WM_TRY(dmalloc(ctx, &ctx->cg_p[1], 3 * n));
  WM_TRY(dmalloc(ctx, &ctx->cg_s, 4));
  WM_TRY(dmalloc(ctx, &ctx->red, wm_cg_partials(g.nx, g.ny)));
  WM_TRY(dmalloc(ctx, &ctx->red_cnt, 8));
  ctx->dep_part_elems = ctx->form == WM_FORM_FP32 ? wm_deposit_part_elems32(g.nx, g.ny)
                                                  : wm_deposit_part_elems(g.nx, g.ny);
  WM_TRY(dmalloc(ctx, &ctx->dep_part, ctx->dep_part_elems));
  WM_TRY(dmalloc(ctx, &ctx->diag, (size_t)nseg * 8 + 2 * (size_t)g.ny + 64));
  WM_TRY(dmalloc(ctx, &ctx->derr, 1));
  WM_TRY(dmalloc(ctx, &ctx->cnt_scratch, nseg));
  WM_TRY(dmalloc(ctx, &ctx->fscratch, 6 * fg.ps));

  WM_HIP(hipMemsetAsync(ctx->red_cnt, 0, 8 * sizeof(unsigned), ctx->stream));
  WM_HIP(hipMemsetAsync(ctx->derr, 0, sizeof(int), ctx->stream));
  WM_HIP(hipMemsetAsync(ctx->np2, 0, sizeof(int) * nseg, ctx->stream));
  WM_HIP(hipMemsetAsync(ctx->cumcnt, 0, sizeof(int) * (size_t)nseg * (g.nx + 1), ctx->stream));
  // field.f90:105-119: df, gkl, uj start at zero
  WM_TRY(fill_zero(ctx, ctx->uf, 6 * fg.ps));
  WM_TRY(fill_zero(ctx, ctx->df, 6 * fg.ps));
  ctx->x0_fresh = false;
  WM_TRY(fill_zero(ctx, ctx->uj, 3 * fg.ps));
  WM_TRY(fill_zero(ctx, ctx->gkl, 3 * n));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  ctx->alloc_done = true;
  return WM_OK;
}

int wm_fetch_derr(wm_ctx *ctx, int *out) {
  int *hp = (int *)ctx->h_pin;
  WM_HIP(hipMemcpyAsync(hp, ctx->derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  *out = *hp;
  if (*out) {
    WM_HIP(hipMemsetAsync(ctx->derr, 0, sizeof(int), ctx->stream));
    WM_HIP(hipStreamSynchronize(ctx->stream));
  }
  return WM_OK;
}

int wm_report_derr(wm_ctx *ctx, int h, const char *where) {
  if (!h) return WM_OK;
  if (h & (DERR_NP_OVER | DERR_MOVER_OVER)) {
    wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np) in %s (%s)", where,
                 (h & DERR_NP_OVER) ? "row capacity" : "migration buffer");
    return WM_ERR_NP_OVER;
  }
  wm_set_error(ctx, WM_ERR_RANGE, "%s: particle outside the cells the reference indexes (%s)",
               where, (h & DERR_JUMP) ? "moved more than one cell" : "cell index out of range");
  return WM_ERR_RANGE;
}

int wm_check_derr(wm_ctx *ctx, const char *where) {
  int h = 0;
  WM_TRY(wm_fetch_derr(ctx, &h));
  return wm_report_derr(ctx, h, where);
}

// ================================================================== C ABI
extern "C" {

int wm_create(wm_ctx **out, int device) {
  if (!out) return WM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
    return WM_ERR_NODEV;
  wm_ctx *ctx = new wm_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipStreamCreate(&ctx->stream2) != hipSuccess) {
    delete ctx;
    return WM_ERR_NODEV;
  }
  // small device-to-host read-backs (solver scalars, error flags) land in
  // pinned memory: a copy to pageable memory is staged by the runtime
  if (hipHostMalloc(&ctx->h_pin, WM_PIN_BYTES, hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return WM_ERR_NODEV;
  }
  for (int k = 0; k < 32; ++k) (void)hipEventCreate(&ctx->ev[k]);
  *out = ctx;
  return WM_OK;
}

int wm_destroy(wm_ctx *ctx) {
  if (!ctx) return WM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_bank(ctx->bank[0]);
  free_bank(ctx->bank[1]);
  free_bank(ctx->mv_up);
  free_bank(ctx->mv_dn);
  if (ctx->tmp32) (void)hipFree(ctx->tmp32);
  void *ptrs[] = {ctx->id[0], ctx->id[1], ctx->np2, ctx->cumcnt, ctx->hole, ctx->nhole,
                  ctx->nup, ctx->ndn, ctx->mvid_up, ctx->mvid_dn, ctx->uf, ctx->df, ctx->uj,
                  ctx->tmp, ctx->gkl, ctx->cg_r, ctx->cg_p[0], ctx->cg_p[1], ctx->cg_s, ctx->red,
                  ctx->red_cnt, ctx->dep_part, ctx->diag, ctx->derr, ctx->stage,
                  ctx->cnt_scratch, ctx->fscratch, ctx->id_bins, ctx->bin_cnt, ctx->bin_sum[0],
                  ctx->bin_sum[1], ctx->mv[0].rec, ctx->mv[1].rec, ctx->mv[0].perm, ctx->mv[1].perm, ctx->cg_z, ctx->cg_q, ctx->cg_x0,
                  ctx->pc_sc, ctx->fcg_s, ctx->mom, ctx->mom_part, ctx->bins.q, ctx->bins.uz,
                  ctx->bstart, ctx->pcnt, ctx->mfill[0], ctx->mfill[1], ctx->rowpend,
                  ctx->fstat, ctx->relay_ctl, ctx->relay_blocks, ctx->blkflag, ctx->relay_segs,
                  ctx->segflag, ctx->rws.q, ctx->rws.uz, ctx->rws_id, ctx->rws_bs,
                  ctx->spill_l, ctx->spill_p, ctx->spill_q, ctx->sfill, ctx->ytmask,
                  ctx->ytpre, ctx->yhist};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  wm_pc_free(ctx);
  for (int k = 0; k < 32; ++k) (void)hipEventDestroy(ctx->ev[k]);
  (void)hipStreamDestroy(ctx->stream);
  (void)hipStreamDestroy(ctx->stream2);
  if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
  delete ctx;
  return WM_OK;
}

const char *wm_last_error(const wm_ctx *ctx) { return ctx ? ctx->msg : "no context"; }

int wm_set_capacity(wm_ctx *ctx, int cap_per_row, int mover_cap) {
  if (!ctx) return WM_ERR_ARG;
  if (ctx->alloc_done) {
    wm_set_error(ctx, WM_ERR_ARG, "capacity must be set before the first device call");
    return WM_ERR_ARG;
  }
  if (cap_per_row < 0 || mover_cap < 0) return WM_ERR_ARG;
  ctx->cap = cap_per_row;
  ctx->mcap = mover_cap;
  return WM_OK;
}

int wm_set_form(wm_ctx *ctx, int form) {
  if (!ctx || (form != WM_FORM_FP64 && form != WM_FORM_FP32)) return WM_ERR_ARG;
  if (ctx->alloc_done) {
    wm_set_error(ctx, WM_ERR_ARG, "the particle form must be set before the first device call");
    return WM_ERR_ARG;
  }
  ctx->form = form;
  return WM_OK;
}

int wm_set_cg_tolerance(wm_ctx *ctx, double err) {
  if (!ctx || !(err > 0.0)) return WM_ERR_ARG;
  ctx->cg_err = err;
  return WM_OK;
}

int wm_mpi_set_init(int nygs, int nyge, int nproc, int *nys, int *nye, int *nup, int *ndown,
                    int *nrank, int *nsize) {
  // mpi_set.f90:19-49 with nsize = 1 (one device holds the whole domain)
  if (nproc != 1) return WM_ERR_ARG;   // mpi_set.f90:30-34 aborts on a mismatch
  const int iwork1 = (nyge - nygs + 1) / 1;
  *nrank = 0;
  *nsize = 1;
  *nys = 0 * iwork1 + nygs;
  *nye = *nys + iwork1 - 1;
  *nup = 0;     // nrank == nsize-1 -> periodic neighbour is rank 0
  *ndown = 0;   // nrank == 0       -> nsize-1 = 0
  return WM_OK;
}

int wm_particle_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs,
                     int nyge, int nys, int nye, double delx, double delt, double c,
                     const double *q, const double *r) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(check_grid(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r, true));
  ctx->pinit = true;
  return WM_OK;
}

int wm_field_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                  int nys, int nye, int mnpr, int ncomw, int opsum, int nerr, double delx,
                  double delt, double c, const double *q, const double *r, double gfac) {
  (void)mnpr; (void)ncomw; (void)opsum; (void)nerr;
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(check_grid(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r, true));
  ctx->g.gfac = gfac;
  ctx->finit = true;
  return WM_OK;
}

int wm_sort_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                 int nys, int nye) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(check_grid(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 1.0, 1.0, 1.0, nullptr,
                    nullptr, false));
  if (ctx->g.nx > 16384) {
    wm_set_error(ctx, WM_ERR_ARG, "sort: nx=%d exceeds the 16384 cells one LDS histogram holds",
                 ctx->g.nx);
    return WM_ERR_ARG;
  }
  ctx->sinit = true;
  return WM_OK;
}

namespace {
// boundary_*__init of one family (boundary_periodic.f90:24-58,
// boundary_shock.f90:26-58, boundary_reconnection.f90:26-58): the same
// arguments; the family picks the device's field, current, potential and
// particle boundaries.
int bc_init(wm_ctx *ctx, int family, int ndim, int np, int nsp, int nxgs, int nxge, int nygs,
            int nyge, int nys, int nye, int nup, int ndown, double delx, double delt, double c) {
  if (!ctx) return WM_ERR_ARG;
  if (nup != 0 || ndown != 0) {
    wm_set_error(ctx, WM_ERR_ARG, "one device context is one periodic rank (nup = ndown = 0)");
    return WM_ERR_ARG;
  }
  const double *q = ctx->have_grid ? ctx->g.q : nullptr;
  const double *r = ctx->have_grid ? ctx->g.r : nullptr;
  WM_TRY(check_grid(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r,
                    q != nullptr));
  if (!q) {
    ctx->g.delx = delx;
    ctx->g.delt = delt;
    ctx->g.c = c;
  }
  if (ctx->binit && ctx->bcf != family) {
    wm_set_error(ctx, WM_ERR_ARG, "a context takes one boundary family (%d, then %d)", ctx->bcf,
                 family);
    return WM_ERR_ARG;
  }
  ctx->bcf = family;
  ctx->binit = true;
  return WM_OK;
}
}  // namespace

int wm_bcp_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                int nys, int nye, int nup, int ndown, int mnpi, int mnpr, int ncomw, int nerr,
                const int *nstat, int nstat_len, double delx, double delt, double c) {
  (void)mnpi; (void)mnpr; (void)ncomw; (void)nerr; (void)nstat; (void)nstat_len;
  return bc_init(ctx, WM_BC_PERIODIC, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup, ndown,
                 delx, delt, c);
}
int wm_bcs_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                int nys, int nye, int nup, int ndown, int mnpi, int mnpr, int ncomw, int nerr,
                const int *nstat, int nstat_len, double delx, double delt, double c) {
  (void)mnpi; (void)mnpr; (void)ncomw; (void)nerr; (void)nstat; (void)nstat_len;
  return bc_init(ctx, WM_BC_SHOCK, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup, ndown,
                 delx, delt, c);
}
int wm_bcr_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                int nys, int nye, int nup, int ndown, int mnpi, int mnpr, int ncomw, int nerr,
                const int *nstat, int nstat_len, double delx, double delt, double c) {
  (void)mnpi; (void)mnpr; (void)ncomw; (void)nerr; (void)nstat; (void)nstat_len;
  return bc_init(ctx, WM_BC_RECONNECTION, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nup,
                 ndown, delx, delt, c);
}

// ------------------------------------------------- host-array entry points
int wm_particle_solv(wm_ctx *ctx, double *gp, const double *up, const double *uf,
                     const int *cumcnt, int nxs, int nxe) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->pinit, "particle__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  WM_TRY(require_fp64(ctx, "particle__solv"));
  WM_TRY(require_no_resident(ctx, "particle__solv"));
  WM_TRY(wm_alloc_device(ctx));
  const std::vector<int> cnt = counts_from_cumcnt(ctx, cumcnt);
  WM_TRY(upload_particles(ctx, up, cnt.data(), 0));
  WM_TRY(upload_cumcnt(ctx, cumcnt));
  WM_TRY(upload_field(ctx, uf, 6, ctx->uf));
  WM_TRY(wm_k_cell_fields(ctx));
  WM_TRY(wm_k_push(ctx, 0, 1));
  WM_TRY(wm_check_derr(ctx, "particle__solv"));
  WM_TRY(download_particles(ctx, gp, cnt.data(), 1));
  // gp(6,:,:,:) = up(6,:,:,:) over the whole capacity (particle.f90:171-175)
  if (ctx->g.ndim == 6) {
    const size_t n = (size_t)ctx->g.np * ctx->g.ny * ctx->g.nsp;
    for (size_t t = 0; t < n; ++t) gp[6 * t + 5] = up[6 * t + 5];
  }
  return WM_OK;
}

int wm_field_fdtd_i(wm_ctx *ctx, double *uf, const double *up, const double *gp,
                    const int *cumcnt, int nxs, int nxe, int bc_family) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->finit, "field__init"));
  static const char *const init_of[3] = {"boundary_periodic__init", "boundary_shock__init",
                                         "boundary_reconnection__init"};
  if (bc_family < 0) bc_family = ctx->binit ? ctx->bcf : WM_BC_PERIODIC;   // the initialised one
  if (bc_family > WM_BC_RECONNECTION) {
    wm_set_error(ctx, WM_ERR_ARG, "boundary family %d is not available", bc_family);
    return WM_ERR_ARG;
  }
  WM_TRY(require(ctx, ctx->binit && ctx->bcf == bc_family, init_of[bc_family]));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  WM_TRY(require_fp64(ctx, "field__fdtd_i"));
  WM_TRY(require_no_resident(ctx, "field__fdtd_i"));
  WM_TRY(wm_alloc_device(ctx));
  const std::vector<int> cnt = counts_from_cumcnt(ctx, cumcnt);
  WM_TRY(upload_particles(ctx, up, cnt.data(), 0));
  WM_TRY(upload_particles(ctx, gp, cnt.data(), 1));
  WM_TRY(upload_cumcnt(ctx, cumcnt));
  WM_TRY(upload_field(ctx, uf, 6, ctx->uf));
  WM_TRY(wm_k_deposit(ctx, 0, 1));
  WM_TRY(wm_k_deposit_gather(ctx));
  WM_TRY(wm_k_bc_curre(ctx, ctx->uj));
  WM_TRY(wm_check_derr(ctx, "field__fdtd_i"));
  WM_TRY(wm_k_field_solve(ctx));
  WM_TRY(download_field(ctx, uf, 6, ctx->uf));
  return WM_OK;
}

int wm_bcp_particle_x(wm_ctx *ctx, double *up, const int *np2) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit && ctx->bcf == WM_BC_PERIODIC, "boundary_periodic__init"));
  WM_TRY(require_fp64(ctx, "boundary_periodic__particle_x"));
  WM_TRY(require_no_resident(ctx, "boundary_periodic__particle_x"));
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_particles(ctx, up, np2, 0));
  WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg_of(ctx), hipMemcpyHostToDevice,
                        ctx->stream));
  WM_TRY(wm_k_particle_x(ctx, 0));
  WM_TRY(wm_check_derr(ctx, "boundary_periodic__particle_x"));
  return download_particles(ctx, up, np2, 0);
}

int wm_bcp_particle_y(wm_ctx *ctx, double *up, int *np2) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit, "boundary_periodic__init"));
  WM_TRY(require_fp64(ctx, "boundary_periodic__particle_y"));
  WM_TRY(require_no_resident(ctx, "boundary_periodic__particle_y"));
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_particles(ctx, up, np2, 0));
  const int nseg = nseg_of(ctx);
  WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_particle_y(ctx, 0, 0));
  WM_TRY(wm_check_derr(ctx, "boundary_periodic__particle_y"));
  std::vector<int> n2(nseg);
  WM_HIP(hipMemcpyAsync(n2.data(), ctx->np2, sizeof(int) * nseg, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < nseg; ++s)
    if (n2[s] > ctx->g.np) {
      wm_set_error(ctx, WM_ERR_NP_OVER, "memory over (np2 > np) %d %d %d %d", ctx->g.np, n2[s],
                   ctx->g.nys + s % ctx->g.ny, 1 + s / ctx->g.ny);
      return WM_ERR_NP_OVER;
    }
  WM_TRY(download_particles(ctx, up, n2.data(), 0));
  memcpy(np2, n2.data(), sizeof(int) * nseg);
  return WM_OK;
}

int wm_sort_bucket(wm_ctx *ctx, double *gp, const double *up, int *cumcnt, const int *np2, int nxs,
                   int nxe) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->sinit, "sort__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  WM_TRY(require_fp64(ctx, "sort__bucket"));
  WM_TRY(require_no_resident(ctx, "sort__bucket"));
  WM_TRY(wm_alloc_device(ctx));
  const int nseg = nseg_of(ctx);
  WM_TRY(upload_particles(ctx, up, np2, 1));
  WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_sort(ctx, 1, 0));
  WM_TRY(wm_check_derr(ctx, "sort__bucket"));
  const int w = ctx->g.nx + 1;
  std::vector<int> cc((size_t)nseg * w);
  WM_HIP(hipMemcpyAsync(cc.data(), ctx->cumcnt, sizeof(int) * cc.size(), hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  // sort.f90:66-69 writes cumcnt(nxs:nxe+1) only
  const int a = nxs - ctx->g.nxgs, b = nxe + 1 - ctx->g.nxgs;
  for (int sg = 0; sg < nseg; ++sg)
    memcpy(cumcnt + (size_t)sg * w + a, cc.data() + (size_t)sg * w + a, sizeof(int) * (b - a + 1));
  return download_particles(ctx, gp, np2, 0);
}

int wm_bcp_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye, int nxgs,
                  int nxge) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit && ctx->bcf == WM_BC_PERIODIC, "boundary_periodic__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  if (nys != ctx->g.nys || nye != ctx->g.nye || nxgs != ctx->g.nxgs || nxge != ctx->g.nxge)
    return WM_ERR_ARG;
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_field(ctx, df, 6, ctx->fscratch));
  WM_TRY(wm_k_bc_dfield(ctx, ctx->fscratch));
  return download_field(ctx, df, 6, ctx->fscratch);
}

int wm_bcp_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye, int nxgs,
                 int nxge) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit && ctx->bcf == WM_BC_PERIODIC, "boundary_periodic__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  if (nys != ctx->g.nys || nye != ctx->g.nye || nxgs != ctx->g.nxgs || nxge != ctx->g.nxge)
    return WM_ERR_ARG;
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_field(ctx, uj, 3, ctx->fscratch));
  WM_TRY(wm_k_bc_curre(ctx, ctx->fscratch));
  return download_field(ctx, uj, 3, ctx->fscratch);
}

namespace {
const char *bc_init_name(int family) {
  return family == WM_BC_SHOCK ? "boundary_shock__init"
         : family == WM_BC_RECONNECTION ? "boundary_reconnection__init"
                                        : "boundary_periodic__init";
}
int require_bc(wm_ctx *ctx, int family) {
  return require(ctx, ctx->binit && ctx->bcf == family, bc_init_name(family));
}

int bc_dfield(wm_ctx *ctx, int family, double *df, int nxs, int nxe, int nys, int nye, int nxgs,
              int nxge) {
  if (!ctx || !df) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, family));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  if (nys != ctx->g.nys || nye != ctx->g.nye || nxgs != ctx->g.nxgs || nxge != ctx->g.nxge)
    return WM_ERR_ARG;
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_field(ctx, df, 6, ctx->fscratch));
  WM_TRY(wm_k_bc_dfield(ctx, ctx->fscratch));
  return download_field(ctx, df, 6, ctx->fscratch);
}

int bc_curre(wm_ctx *ctx, int family, double *uj, int nxs, int nxe, int nys, int nye, int nxgs,
             int nxge) {
  if (!ctx || !uj) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, family));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  if (nys != ctx->g.nys || nye != ctx->g.nye || nxgs != ctx->g.nxgs || nxge != ctx->g.nxge)
    return WM_ERR_ARG;
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_field(ctx, uj, 3, ctx->fscratch));
  WM_TRY(wm_k_bc_curre(ctx, ctx->fscratch));
  return download_field(ctx, uj, 3, ctx->fscratch);
}

// phi(nxs-1:nxe+1, nys-1:nye+1) of component l
int bc_phi(wm_ctx *ctx, int family, double *phi, int nxs, int nxe, int nys, int nye, int l) {
  if (!ctx || !phi) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, family));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  if (nys != ctx->g.nys || nye != ctx->g.nye || l < 1 || l > 3) return WM_ERR_ARG;
  WM_TRY(wm_alloc_device(ctx));
  const size_t n = (size_t)(nxe - nxs + 3) * (ctx->g.ny + 2);
  WM_TRY(ensure_stage(ctx, n));
  WM_HIP(hipMemcpyAsync(ctx->stage, phi, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_bc_phi(ctx, ctx->stage, l));
  WM_HIP(hipMemcpyAsync(phi, ctx->stage, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

// The walls' particle_x (family shock or reconnection) or the shock's
// injection edge (inj) on host arrays.
int bc_walls_x(wm_ctx *ctx, int family, bool inj, double *up, const int *np2, int nxs, int nxe,
               double u0, const char *what) {
  if (!ctx || !up || !np2) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, family));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  WM_TRY(require_fp64(ctx, what));
  WM_TRY(require_no_resident(ctx, what));
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(upload_particles(ctx, up, np2, 0));
  WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg_of(ctx), hipMemcpyHostToDevice,
                        ctx->stream));
  WM_TRY(inj ? wm_k_injection(ctx, 0, u0) : wm_k_particle_x(ctx, 0));
  WM_TRY(wm_check_derr(ctx, what));
  return download_particles(ctx, up, np2, 0);
}

int bc_mom(wm_ctx *ctx, int family, double *mom) {
  if (!ctx || !mom) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, family));
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(wm_mom_buffers(ctx));
  const size_t n = wm_mom_elems(ctx);
  WM_HIP(hipMemcpyAsync(ctx->mom, mom, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_bc_mom(ctx, ctx->mom));
  WM_HIP(hipMemcpyAsync(mom, ctx->mom, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}
}  // namespace

int wm_bcp_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye, int l) {
  return bc_phi(ctx, WM_BC_PERIODIC, phi, nxs, nxe, nys, nye, l);
}

// ---- boundary_shock__* (proj/shock/boundary_shock.f90)
int wm_bcs_particle_x(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe) {
  return bc_walls_x(ctx, WM_BC_SHOCK, false, up, np2, nxs, nxe, 0.0, "boundary_shock__particle_x");
}
int wm_bcs_injection(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe, double u0) {
  return bc_walls_x(ctx, WM_BC_SHOCK, true, up, np2, nxs, nxe, u0, "boundary_shock__injection");
}
int wm_bcs_particle_y(wm_ctx *ctx, double *up, int *np2) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, WM_BC_SHOCK));
  return wm_bcp_particle_y(ctx, up, np2);
}
int wm_bcs_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye, int nxgs, int nxge) {
  return bc_dfield(ctx, WM_BC_SHOCK, df, nxs, nxe, nys, nye, nxgs, nxge);
}
int wm_bcs_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye, int nxgs, int nxge) {
  return bc_curre(ctx, WM_BC_SHOCK, uj, nxs, nxe, nys, nye, nxgs, nxge);
}
int wm_bcs_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye, int l) {
  return bc_phi(ctx, WM_BC_SHOCK, phi, nxs, nxe, nys, nye, l);
}
int wm_bcs_mom(wm_ctx *ctx, double *mom) { return bc_mom(ctx, WM_BC_SHOCK, mom); }

// ---- boundary_reconnection__* (proj/reconnection/boundary_reconnection.f90)
int wm_bcr_particle_x(wm_ctx *ctx, double *up, const int *np2, int nxs, int nxe) {
  return bc_walls_x(ctx, WM_BC_RECONNECTION, false, up, np2, nxs, nxe, 0.0,
                    "boundary_reconnection__particle_x");
}
int wm_bcr_particle_y(wm_ctx *ctx, double *up, int *np2) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require_bc(ctx, WM_BC_RECONNECTION));
  return wm_bcp_particle_y(ctx, up, np2);
}
int wm_bcr_dfield(wm_ctx *ctx, double *df, int nxs, int nxe, int nys, int nye, int nxgs, int nxge) {
  return bc_dfield(ctx, WM_BC_RECONNECTION, df, nxs, nxe, nys, nye, nxgs, nxge);
}
int wm_bcr_curre(wm_ctx *ctx, double *uj, int nxs, int nxe, int nys, int nye, int nxgs, int nxge) {
  return bc_curre(ctx, WM_BC_RECONNECTION, uj, nxs, nxe, nys, nye, nxgs, nxge);
}
int wm_bcr_phi(wm_ctx *ctx, double *phi, int nxs, int nxe, int nys, int nye, int l) {
  return bc_phi(ctx, WM_BC_RECONNECTION, phi, nxs, nxe, nys, nye, l);
}
int wm_bcr_mom(wm_ctx *ctx, double *mom) { return bc_mom(ctx, WM_BC_RECONNECTION, mom); }

// Device-resident runs of a wall family: the active columns and the shock's
// upstream four-velocity (the arguments the reference passes each step).
int wm_set_walls(wm_ctx *ctx, int nxs, int nxe, double u0) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit && ctx->bcf != WM_BC_PERIODIC,
                 "boundary_shock__init or boundary_reconnection__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  ctx->u0 = u0;
  return WM_OK;
}

// ------------------------------------------------------ resident state
int wm_state_upload(wm_ctx *ctx, const double *up, const int *np2, const int *cumcnt,
                    const double *uf) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->pinit, "particle__init"));
  WM_TRY(wm_alloc_device(ctx));
  const int nseg = nseg_of(ctx);
  if (ctx->form == WM_FORM_FP32) {
    // the rows go straight into the cell bins (cumcnt only checked)
    WM_TRY(upload_cumcnt(ctx, cumcnt));
    WM_TRY(fresh_bins(ctx));
    WM_TRY(upload_bins(ctx, up, np2));
  } else {
    WM_TRY(upload_particles(ctx, up, np2, 0));
    WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
    WM_TRY(upload_cumcnt(ctx, cumcnt));
  }
  WM_TRY(upload_field(ctx, uf, 6, ctx->uf));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  WM_TRY(wm_check_derr(ctx, "state upload"));
  ctx->resident = true;
  return WM_OK;
}

int wm_state_download(wm_ctx *ctx, double *up, int *np2, int *cumcnt, double *uf) {
  if (!ctx) return WM_ERR_ARG;
  if (!ctx->alloc_done) return WM_ERR_NOT_INIT;
  const int nseg = nseg_of(ctx);
  const bool f32 = ctx->form == WM_FORM_FP32;
  // FP32: the bins' counts give np2 and cumcnt (sorted by cell, as sort.f90
  // leaves them)
  if (f32 && (up || np2 || cumcnt)) {
    WM_TRY(wm_settle_movers(ctx));
    WM_TRY(wm_k_bin_rows(ctx));
  }
  std::vector<int> n2(nseg);
  WM_HIP(hipMemcpyAsync(n2.data(), ctx->np2, sizeof(int) * nseg, hipMemcpyDeviceToHost, ctx->stream));
  if (cumcnt)
    WM_HIP(hipMemcpyAsync(cumcnt, ctx->cumcnt, sizeof(int) * (size_t)nseg * (ctx->g.nx + 1),
                          hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  if (f32) WM_TRY(wm_check_derr(ctx, "state download"));
  if (np2) memcpy(np2, n2.data(), sizeof(int) * nseg);
  if (up) WM_TRY(f32 ? download_bins(ctx, up, n2.data()) : download_particles(ctx, up, n2.data(), 0));
  if (f32 && up) WM_TRY(wm_check_derr(ctx, "state download"));
  if (uf) WM_TRY(download_field(ctx, uf, 6, ctx->uf));
  return WM_OK;
}

int wm_state_init_weibel(wm_ctx *ctx, uint64_t seed, int ppc, double vth, double udrift) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->pinit, "particle__init"));
  WM_TRY(require(ctx, ctx->sinit, "sort__init"));
  WM_TRY(wm_alloc_device(ctx));
  if (ctx->form == WM_FORM_FP32) {
    // straight into the cell bins, stable by cell within each row
    WM_TRY(fresh_bins(ctx));
    WM_TRY(wm_k_init_weibel(ctx, seed, ppc, vth, udrift));
  } else {
    WM_TRY(wm_k_init_weibel(ctx, seed, ppc, vth, udrift));
    // stable by cell within each row: the order of sort__bucket
    WM_TRY(wm_k_sort(ctx, 1, 0));
  }
  WM_TRY(fill_zero(ctx, ctx->uf, 6 * ctx->fg.ps));
  WM_TRY(wm_check_derr(ctx, "initial state"));
  ctx->resident = true;
  return WM_OK;
}

namespace {
// Routes the launches of a scope to another stream of the context.
struct StreamSwap {
  wm_ctx *c;
  hipStream_t saved;
  StreamSwap(wm_ctx *ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { ctx->stream = s; }
  ~StreamSwap() { c->stream = saved; }
};
}  // namespace

int wm_step(wm_ctx *ctx, int nsteps) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->pinit, "particle__init"));
  WM_TRY(require(ctx, ctx->finit, "field__init"));
  WM_TRY(require(ctx, ctx->sinit, "sort__init"));
  WM_TRY(require(ctx, ctx->binit, "boundary_periodic__init"));
  if (ctx->bcf != WM_BC_PERIODIC && ctx->form == WM_FORM_FP32) {
    wm_set_error(ctx, WM_ERR_ARG, "the wall boundary families run in the FP64 particle form");
    return WM_ERR_ARG;
  }
  WM_TRY(wm_alloc_device(ctx));
  double acc[WM_NPHASE] = {0};
  for (int it = 0; it < nsteps; ++it) {
    hipEvent_t *e = ctx->ev;
    WM_HIP(hipEventRecord(e[0], ctx->stream));
    // particle__solv + ele_cur (proj/weibel/app.f90:102, field.f90:121)
    const bool f32 = ctx->form == WM_FORM_FP32;
    WM_TRY(f32 ? wm_k_cell_fields32(ctx) : wm_k_cell_fields(ctx));
    WM_HIP(hipEventRecord(e[1], ctx->stream));
    if (f32) {
      // Cell bins: the push wraps x itself, and moves the last push's movers
      // into their bins before it pushes (the particle boundaries and the
      // sort of app.f90:104-106 reduce to that).  Beside the solve, on the
      // second stream, the fill check lists the blocks of cells whose bins
      // would overflow with this push's movers and they are laid out anew.
      WM_TRY(wm_k_push_deposit32(ctx));
      WM_HIP(hipEventRecord(e[2], ctx->stream));
      WM_HIP(hipStreamWaitEvent(ctx->stream2, e[2], 0));
      StreamSwap on2(ctx, ctx->stream2);
      WM_HIP(hipEventRecord(e[5], ctx->stream));
      WM_TRY(wm_k_fill_check(ctx));
      WM_HIP(hipEventRecord(e[6], ctx->stream));
      WM_TRY(wm_relayout(ctx));
      WM_HIP(hipEventRecord(e[7], ctx->stream));
    } else {
      // The particle boundaries and the sort (app.f90:104-106) read only the
      // pushed particles, and the field solve (app.f90:103) only the current
      // and the fields: they run concurrently, the particles on the second
      // stream, and the next step waits for both.
      WM_TRY(wm_k_push_deposit(ctx, 0, 1));
      WM_HIP(hipEventRecord(e[2], ctx->stream));
      WM_HIP(hipStreamWaitEvent(ctx->stream2, e[2], 0));
      StreamSwap on2(ctx, ctx->stream2);
      WM_HIP(hipEventRecord(e[5], ctx->stream));
      {
        // app.f90:104-105 (x fused into y pass 1): the periodic wrap, the
        // reconnection's walls (proj/reconnection/app.f90:104), or none
        // (the shock's edges came with the push, proj/shock/app.f90:113)
        const int xmode = ctx->bcf == WM_BC_PERIODIC ? 1 : ctx->bcf == WM_BC_RECONNECTION ? 2 : 0;
        // app.f90:105-106: pass 1, then pass 2 fused with the sort
        WM_TRY(wm_k_particle_y_sort(ctx, 1, 0, xmode, e[6]));
      }
      WM_HIP(hipEventRecord(e[7], ctx->stream));
    }
    WM_TRY(f32 ? wm_k_deposit_gather32(ctx) : wm_k_deposit_gather(ctx));
    WM_TRY(wm_k_bc_curre(ctx, ctx->uj));          // field.f90:122
    WM_HIP(hipEventRecord(e[3], ctx->stream));
    WM_TRY(wm_k_field_solve(ctx));                // field.f90:125-184
    WM_HIP(hipEventRecord(e[4], ctx->stream));
    WM_HIP(hipStreamWaitEvent(ctx->stream, e[7], 0));
    WM_HIP(hipEventRecord(e[8], ctx->stream));
    WM_HIP(hipEventSynchronize(e[8]));
    // phases: [0..3] on the first stream (e0..e4), [4] particle boundaries
    // (e5..e6) and [5] sort (e6..e7) on the second (FP64; in the FP32 form
    // the push holds them, and [4] is the fill check, [5] the re-layout)
    const int from[WM_NPHASE] = {0, 1, 2, 3, 5, 6}, to[WM_NPHASE] = {1, 2, 3, 4, 6, 7};
    for (int k = 0; k < WM_NPHASE; ++k) {
      float ms = 0.f;
      WM_HIP(hipEventElapsedTime(&ms, e[from[k]], e[to[k]]));
      acc[k] += ms;
    }
    WM_TRY(wm_check_derr(ctx, "step"));
  }
  for (int k = 0; k < WM_NPHASE; ++k) ctx->phase_ms[k] = acc[k];
  return WM_OK;
}

int wm_energy(wm_ctx *ctx, double *out) {
  if (!ctx || !out) return WM_ERR_ARG;
  if (!ctx->alloc_done) return WM_ERR_NOT_INIT;
  return wm_k_energy(ctx, 0, out);
}

// ------------------------------------------------------ moment diagnostics
int wm_mom_calc_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs,
                     int nyge, int nys, int nye, double delx, double delt, double c,
                     const double *q, const double *r) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(check_grid(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r, true));
  if (ctx->g.nx > 16384) {     // mom_calc__nvt's host-array form bins rows with the sort
    wm_set_error(ctx, WM_ERR_ARG, "mom_calc: nx=%d exceeds the 16384 cells one LDS histogram holds",
                 ctx->g.nx);
    return WM_ERR_ARG;
  }
  ctx->minit = true;
  return WM_OK;
}

int wm_mom_calc_accl(wm_ctx *ctx, double *gp, const double *up, const double *uf,
                     const int *cumcnt, int nxs, int nxe) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->minit, "mom_calc__init"));
  WM_TRY(check_bounds(ctx, nxs, nxe));
  WM_TRY(require_fp64(ctx, "mom_calc__accl"));
  WM_TRY(require_no_resident(ctx, "mom_calc__accl"));
  WM_TRY(wm_alloc_device(ctx));
  const std::vector<int> cnt = counts_from_cumcnt(ctx, cumcnt);
  WM_TRY(upload_particles(ctx, up, cnt.data(), 0));
  WM_TRY(upload_cumcnt(ctx, cumcnt));
  WM_TRY(upload_field(ctx, uf, 6, ctx->uf));
  WM_TRY(wm_k_accl(ctx, 0, 1));
  WM_TRY(wm_check_derr(ctx, "mom_calc__accl"));
  // gp(1:5) of the particles in cumcnt; gp(6) is not written (mom_calc.f90:151-155)
  return download_particles(ctx, gp, cnt.data(), 1, 5);
}

int wm_mom_calc_nvt(wm_ctx *ctx, double *mom, const double *up, const int *np2) {
  if (!ctx || !mom) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->minit, "mom_calc__init"));
  WM_TRY(require_fp64(ctx, "mom_calc__nvt"));
  WM_TRY(require_no_resident(ctx, "mom_calc__nvt"));
  WM_TRY(wm_alloc_device(ctx));
  // the rows' particles, in any order, are binned by cell first (the sums
  // then run per cell; the reference sums in row order)
  const int nseg = nseg_of(ctx);
  WM_TRY(upload_particles(ctx, up, np2, 1));
  WM_HIP(hipMemcpyAsync(ctx->np2, np2, sizeof(int) * nseg, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_sort(ctx, 1, 0));
  WM_TRY(wm_k_moments(ctx, false));
  WM_TRY(wm_check_derr(ctx, "mom_calc__nvt"));
  WM_HIP(hipMemcpyAsync(mom, ctx->mom, sizeof(double) * wm_mom_elems(ctx), hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_bcp_mom(wm_ctx *ctx, double *mom) {
  if (!ctx || !mom) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->binit && ctx->bcf == WM_BC_PERIODIC, "boundary_periodic__init"));
  WM_TRY(wm_alloc_device(ctx));
  WM_TRY(wm_mom_buffers(ctx));
  const size_t n = wm_mom_elems(ctx);
  WM_HIP(hipMemcpyAsync(ctx->mom, mom, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  WM_TRY(wm_k_bc_mom(ctx, ctx->mom));
  WM_HIP(hipMemcpyAsync(mom, ctx->mom, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_moments(wm_ctx *ctx, double *mom) {
  if (!ctx || !mom) return WM_ERR_ARG;
  WM_TRY(require(ctx, ctx->minit, "mom_calc__init"));
  if (!ctx->alloc_done) {
    wm_set_error(ctx, WM_ERR_NOT_INIT, "moments: no device-resident state (wm_state_upload or "
                                       "wm_state_init_weibel first)");
    return WM_ERR_NOT_INIT;
  }
  if (ctx->form == WM_FORM_FP32) WM_TRY(wm_settle_movers(ctx));
  hipEvent_t *e = ctx->ev + 20;
  WM_HIP(hipEventRecord(e[0], ctx->stream));
  WM_TRY(wm_k_moments(ctx, true));
  WM_HIP(hipEventRecord(e[1], ctx->stream));
  WM_TRY(wm_k_bc_mom(ctx, ctx->mom));
  WM_HIP(hipEventRecord(e[2], ctx->stream));
  WM_TRY(wm_check_derr(ctx, "mom_calc__nvt"));
  float ms = 0.f;
  WM_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
  ctx->phase_ms[WM_NPHASE] = ms;
  WM_HIP(hipEventElapsedTime(&ms, e[1], e[2]));
  ctx->phase_ms[WM_NPHASE + 1] = ms;
  WM_HIP(hipMemcpyAsync(mom, ctx->mom, sizeof(double) * wm_mom_elems(ctx), hipMemcpyDeviceToHost,
                        ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  return WM_OK;
}

int wm_cg_iters(wm_ctx *ctx, int *iters3) {
  if (!ctx || !iters3) return WM_ERR_ARG;
  for (int l = 0; l < 3; ++l) iters3[l] = ctx->last_iters[l];
  return WM_OK;
}

int wm_phase_times(wm_ctx *ctx, double *ms, int n) {
  if (!ctx || !ms) return WM_ERR_ARG;
  const int m = std::min(n, (int)WM_NPHASE + 2);
  for (int k = 0; k < m; ++k) ms[k] = ctx->phase_ms[k];
  return m;
}

int wm_particle_count(wm_ctx *ctx, long long *count) {
  if (!ctx || !count) return WM_ERR_ARG;
  if (!ctx->alloc_done) return WM_ERR_NOT_INIT;
  // FP32: the bin counts; FP64: np2
  const bool f32 = ctx->form == WM_FORM_FP32;
  const size_t n = f32 ? (size_t)nseg_of(ctx) * ctx->g.nx : (size_t)nseg_of(ctx);
  std::vector<int> c(n);
  WM_HIP(hipMemcpyAsync(c.data(), f32 ? ctx->bin_cnt : ctx->np2, sizeof(int) * n,
                        hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  long long t = 0;
  for (int v : c) t += v;
  *count = t;
  return WM_OK;
}

int wm_id_checksum(wm_ctx *ctx, unsigned long long *sum_xor) {
  if (!ctx || !sum_xor) return WM_ERR_ARG;
  if (!ctx->alloc_done) return WM_ERR_NOT_INIT;
  if (ctx->form == WM_FORM_FP32) WM_TRY(wm_settle_movers(ctx));
  unsigned long long *d = nullptr;
  WM_HIP(hipMalloc((void **)&d, 2 * sizeof(unsigned long long)));
  int st = WM_OK;
  if (hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), ctx->stream) != hipSuccess)
    st = WM_ERR_HIP;
  if (st == WM_OK) st = wm_k_id_checksum(ctx, d);
  if (st == WM_OK && (hipMemcpyAsync(sum_xor, d, 2 * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                      hipStreamSynchronize(ctx->stream) != hipSuccess))
    st = WM_ERR_HIP;
  (void)hipFree(d);
  return st;
}

int wm_capacity_stats(wm_ctx *ctx, long long *out, int n) {
  if (!ctx || !out) return WM_ERR_ARG;
  if (!ctx->alloc_done || ctx->form != WM_FORM_FP32) {
    wm_set_error(ctx, WM_ERR_ARG, "capacity statistics: no FP32 device state");
    return WM_ERR_ARG;
  }
  int st[WM_NSTAT];
  WM_HIP(hipMemcpyAsync(st, ctx->fstat, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  WM_HIP(hipStreamSynchronize(ctx->stream));
  const long long v[12] = {st[0], st[1], st[2], st[3], st[4], st[6], ctx->cap,
                           ctx->mrow, wm_push32_pool(), ctx->srow, st[5], st[7]};
  const int m = std::min(n, 12);
  for (int k = 0; k < m; ++k) out[k] = v[k];
  return m;
}

// ------------------------------------------- learned preconditioner
namespace {
int require_grid(wm_ctx *ctx) {
  WM_TRY(require(ctx, ctx->finit, "field__init"));
  WM_TRY(require(ctx, ctx->pinit, "particle__init"));
  return wm_alloc_device(ctx);
}
}  // namespace

int wm_set_solver(wm_ctx *ctx, int solver) {
  if (!ctx || (solver != WM_SOLVER_CG && solver != WM_SOLVER_FCG)) return WM_ERR_ARG;
  if (solver == WM_SOLVER_FCG && !wm_pc_ready(ctx)) {
    wm_set_error(ctx, WM_ERR_ARG, "flexible CG needs preconditioner parameters (train or set them)");
    return WM_ERR_ARG;
  }
  ctx->solver = solver;
  return WM_OK;
}

int wm_pc_collect(wm_ctx *ctx, int nslot, int per_solve) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return wm_pc_collect_cfg(ctx, nslot, per_solve);
}

int wm_pc_pair_count(wm_ctx *ctx) { return ctx ? wm_pc_pairs(ctx) : 0; }

int wm_pc_init(wm_ctx *ctx, uint64_t seed) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return wm_pc_init_params(ctx, seed);
}

int wm_pc_train(wm_ctx *ctx, int steps, double lr, uint64_t seed, double *loss2) {
  if (!ctx) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return wm_pc_train_dev(ctx, steps, lr, seed, loss2);
}

int wm_pc_num_params(void) { return wm_pc_nparam(); }

int wm_pc_params(wm_ctx *ctx, float *theta, int n, int set) {
  if (!ctx || !theta) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return set ? wm_pc_set_params_dev(ctx, theta, n) : wm_pc_get_params_dev(ctx, theta, n);
}

int wm_pc_apply(wm_ctx *ctx, const double *r, double *z) {
  if (!ctx || !r || !z) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return wm_pc_apply_host(ctx, r, z);
}

int wm_pc_apply_noalloc(wm_ctx *ctx, const double *r, double *z, int nx, int ny) {
  if (!ctx || !r || !z) return WM_ERR_ARG;
  if (!ctx->pc) {                    // the parameters live in the context: nothing to apply
    wm_set_error(ctx, WM_ERR_ARG, "learned preconditioner: set or train the parameters first");
    return WM_ERR_ARG;
  }
  return wm_pc_apply_noalloc_dev(ctx, r, z, nx, ny);
}

int wm_pc_grad(wm_ctx *ctx, const double *r, const double *e, float *grad, double *loss) {
  if (!ctx || !r || !e || !grad || !loss) return WM_ERR_ARG;
  WM_TRY(require_grid(ctx));
  return wm_pc_grad_host(ctx, r, e, grad, loss);
}

int wm_solve_stats(wm_ctx *ctx, int *fcg3, int *plain3) {
  if (!ctx || !fcg3 || !plain3) return WM_ERR_ARG;
  if (ctx->solver == WM_SOLVER_FCG) {
    wm_pc_get_stats(ctx, fcg3, plain3);
  } else {
    for (int l = 0; l < 3; ++l) {
      fcg3[l] = 0;
      plain3[l] = ctx->last_iters[l];
    }
  }
  return WM_OK;
}

}  // extern "C"
