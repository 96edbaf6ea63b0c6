// wm_internal.h -- device layout, context and shared helpers of libwmpic.
//
// Device layout (MI355X, HBM-resident for the whole run):
//   Particles: structure-of-arrays, one segment per (species s, row j) of
//   `cap` slots, segment index seg = s*ny + (j - nys).  Slot k of a segment is
//   the reference's particle index ii = k+1 of row j (proj/weibel/app.f90:281
//   up(ndim,np,nys:nye,nsp)), so row-local orders, np2 and cumcnt keep the
//   reference's meaning exactly.  Two position/momentum banks: bank 0 plays
//   `up`, bank 1 plays `gp`.  IDs live in their own ping-pong pair.
//   Fields: one plane per component, x fastest, ghost-inclusive
//   (nx+4) x (ny+4) for uf/df/uj (the reference's nxgs-2:nxge+2,
//   nys-2:nye+2 bounds), planes of (nx+2) x (ny+2) for the cell-centred
//   push fields, nx x ny for the CG vectors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/wmpic.h"
#include "wm_records.h"

#define WM_MAXSP 8

// ---------------------------------------------------------------- checking
#define WM_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      wm_set_error(ctx, WM_ERR_HIP, "%s: %s (%s:%d)", #call,                 \
                   hipGetErrorString(e_), __FILE__, __LINE__);               \
      return WM_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

#define WM_TRY(expr)          \
  do {                        \
    int s_ = (expr);          \
    if (s_ != WM_OK) return s_; \
  } while (0)

// Device-side error bits (set with vector atomics, read back by the host).
enum : int {
  DERR_RANGE = 1,     // particle outside the cells the reference indexes
  DERR_NP_OVER = 2,   // a row exceeds the device capacity
  DERR_MOVER_OVER = 4,// a row has more migrating particles than the buffer
  DERR_JUMP = 8,      // a particle moved more than one cell in one step
  DERR_CG_ITEMAX = 16 // CG reached ite_max
};

// ------------------------------------------------------------- parameters
struct wm_grid {
  int ndim, np, nsp;
  int nxgs, nxge, nygs, nyge, nys, nye;
  int nx, ny;
  double delx, delt, c, gfac;
  double q[WM_MAXSP], r[WM_MAXSP];
};

// The x side of the boundaries as the kernels see it: wall = 0 periodic,
// else the family; [xa, xb] the active interior columns, 0-based.
struct XSide {
  int wall, xa, xb;
};

// Field plane geometry (ghost-inclusive, 2 ghost cells).
struct FieldGeom {
  int nx, ny;      // interior
  int gx, gy;      // nx+4, ny+4
  size_t ps;       // plane stride (elements)
  // index of reference (i, j) in a ghost-inclusive plane
  __host__ __device__ inline size_t at(int i, int j, int nxgs, int nys) const {
    return (size_t)(i - (nxgs - 2)) + (size_t)gx * (size_t)(j - (nys - 2));
  }
};

// CG per-component scalar state (device).
#define WM_PIN_BYTES 4096

struct CgScal {
  double sum_g, eps, sumr, sum2, av, bv;
  int active, ite, err, pad;
};

// Flexible-CG per-component scalar state (device).
struct FcgScal {
  double eps, sumr, rz, rz_prev, zq, pq, alpha, beta;
  int active, it, bad, itmax;
};

// FP32 form, cell bins: the movers of one push, one 32-B record each, in a
// pool per row (all species) that the push workgroups fill chunk by chunk.
struct __align__(16) MoverRec {   // 32 B: two 16-byte halves
  float ox1, oy1;              // offsets in the destination cell
  float ux, uy, uz;
  int sd;                      // source slot | direction index << 24 (index k of
                               // direction (dy + 1) * 3 + (dx + 1): 0..3, then 5..8 as 4..7)
  int64_t id;
};
struct Movers {
  MoverRec *rec = nullptr;
  // per bin with more than WM_SUM_MV movers, beside its records: the
  // records' indices (within the bin's) in direction order, so the migration
  // reads the movers of one direction without scanning the bin's list
  int *perm = nullptr;
};
// Per bin after the push: what the migration needs of the movers that left
// it, in one 32-byte load.
#define WM_SUM_MV 8            // movers of a bin listed in its summary
#define WM_SLOT_BITS 13        // source slots below 2^13 (bins of at most 8192 slots)
#define WM_BIN_MAX (1 << WM_SLOT_BITS)
#define WM_BIN_ALIGN 8         // bins start on 8-slot (128-byte float4) boundaries
struct __align__(32) BinSum {
  uint8_t cnt[8];              // movers out per direction index (at most 255 each)
  uint32_t first;              // the first of its movers' records in the row's pool
  uint32_t m;                  // movers out
  uint16_t mv[WM_SUM_MV];      // the first movers in list order: slot | index << 13
};
// Device statistics of the FP32 form's capacities (maxima since the state was
// made): [0] particles in one bin, [1] particles leaving one bin in a push,
// [2] movers of one push workgroup, [3] records of one row's mover pool,
// [4] entries of one row's spill of the push's mover lists, [5] bins found
// full by the fill check, [6] blocks and [7] segments laid out anew (sums).
#define WM_NSTAT 8
#define WM_RB 64               // cells of a re-layout block

using PBank = Rec64;

struct wm_pc;   // learned preconditioner (wm_pc.hip)

struct wm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // particle boundaries + sort, beside the field solve
  int status = WM_OK;
  char msg[512] = {0};

  // init flags and parameters (one grid per context)
  bool pinit = false, finit = false, sinit = false, binit = false, minit = false;
  bool have_grid = false;
  // The boundary family (WM_BC_*) whose __init was called, the active x
  // range [xs, xe] the reference passes as nxs, nxe (xe = 0: nxgs..nxge),
  // and the shock's upstream four-velocity u0 (its injection edge).
  int bcf = 0;
  int xs = 0, xe = 0;
  double u0 = 0.0;
  wm_grid g{};
  double cg_err = 1e-6;    // field.f90:339
  int ite_max = 100;       // field.f90:337

  // capacities
  int cap = 0;             // device slots per (species, row)
  int mcap = 0;            // migrating particles per row and direction
  bool alloc_done = false;
  // a device-resident state exists (wm_state_upload / wm_state_init_weibel):
  // the host-array entry points, which stage their arguments through the
  // same particle banks, fields and counts, refuse from then on
  bool resident = false;

  // particles: FP64 form (bank) or FP32 form (bins), chosen before allocation
  int form = WM_FORM_FP64;
  PBank bank[2]{};
  float *tmp32 = nullptr;  // cell-centred fields of the FP32 push
  Rec64 b64(int k) const { return bank[k]; }
  int64_t *id[2]{};
  int idcur = 0;           // which id array goes with bank 0 / bank 1 data
  int *np2 = nullptr;      // [nseg]
  int *cumcnt = nullptr;   // [nseg*(nx+1)]

  // particle_y workspace
  int *hole = nullptr;     // [nseg*2*mcap]
  int *nhole = nullptr, *nup = nullptr, *ndn = nullptr;  // [nseg]
  PBank mv_up{}, mv_dn{};
  int64_t *mvid_up = nullptr, *mvid_dn = nullptr;
  // the resident step's fused pass 2 + sort: per tile of 64 slots of a row
  // segment, pass 1's leaver mask and the holes before the tile; per
  // segment, the histogram of the cells the row holds after pass 2
  unsigned long long *ytmask = nullptr;   // [nseg*ytiles]
  int *ytpre = nullptr;                   // [nseg*ytiles]
  int *yhist = nullptr;                   // [nseg*nx], zero between steps
  int ytiles = 0;

  // fields
  FieldGeom fg{};
  double *uf = nullptr, *df = nullptr, *uj = nullptr;   // 6, 6, 3 planes
  double *tmp = nullptr;   // 6 planes (nx+2)*(ny+2)
  double *gkl = nullptr;   // 3 planes nx*ny
  double *cg_phi = nullptr, *cg_r = nullptr, *cg_b = nullptr, *cg_ap = nullptr;
  double *cg_p[2]{};
  int cg_pcur = 0;
  CgScal *cg_s = nullptr;  // [3]
  double *red = nullptr;   // reduction partials
  unsigned *red_cnt = nullptr;
  double *dep_part = nullptr; // deposit tiles
  size_t dep_part_elems = 0;
  double *diag = nullptr;  // diagnostics scratch
  int *derr = nullptr;     // device error bits
  void *h_pin = nullptr;   // pinned host scratch for small read-backs (WM_PIN_BYTES)

  // host staging (drop-in host-array entry points)
  double *stage = nullptr; size_t stage_elems = 0;
  int *stage_i = nullptr; size_t stage_ielems = 0;

  // timing (per-phase HIP events of the last wm_step call)
  hipEvent_t ev[32]{};
  bool ev_ready = false;
  int last_iters[3] = {0, 0, 0};
  double phase_ms[8] = {0};
  int *cnt_scratch = nullptr;    // [nseg] counts for layout conversion
  // FP32 form: particles live in cell bins inside the row segments.  Bin
  // (s, jl, i), index seg * nx + i with seg = s * ny + jl, occupies slots
  // [bstart[seg * (nx + 1) + i], bstart[seg * (nx + 1) + i + 1]) of segment seg
  // (seg * cap ..) and holds bin_cnt[bin] particles at its front.  The layout
  // gives each bin room beyond its count; a bin that would overflow gets its
  // block of cells laid out anew (wm_relayout).
  Bins32 bins;
  int *bstart = nullptr;
  int *bin_cnt = nullptr;        // particles per bin
  int64_t *id_bins = nullptr;    // null when ndim = 5 (no IDs)
  // The movers of a push (records in per-row pools, summary per bin) join
  // their destination bins in the next push's prologue, in the settle kernel
  // before the bins are read otherwise, or in a re-layout of their rows; two
  // buffers, because a push reads the last push's movers while it writes
  // its own.
  Movers mv[2]{};
  int *mfill[2] = {nullptr, nullptr};   // records used per row
  int mrow = 0;                  // records per row pool
  // The push lists its movers in LDS; a workgroup with more movers than its
  // LDS list holds continues the list in blocks of its row's spill pool
  // (srow entries per row; blocks used in sfill).
  int *spill_l = nullptr;
  float4 *spill_p = nullptr;
  float *spill_q = nullptr;
  int *sfill = nullptr;
  int srow = 0;
  BinSum *bin_sum[2] = {nullptr, nullptr};
  int mv_last = 0;               // the buffer the last push wrote
  bool mv_pending = false;       // some rows' bins still await its movers
  int *rowpend = nullptr;        // [ny] 1: the row's bins await the last push's movers
  int *pcnt = nullptr;           // [nbin] counts once the pending movers are placed
  int *fstat = nullptr;          // [WM_NSTAT] capacity statistics
  // Re-layout, on the device after the fill check: blocks of WM_RB cells
  // with a bin that would overflow are listed (blkflag dedups), each laid out
  // anew inside its own span; a block whose span cannot hold its bins lists
  // its segment, laid out anew as a whole.  Bins move verbatim (holes
  // included), so the pending movers' bin-relative slots stay valid.
  int *relay_ctl = nullptr;      // [4] listed blocks, listed segments
  int *relay_blocks = nullptr;   // [nseg * nblk] listed blocks (seg * nblk + block)
  int *relay_segs = nullptr;     // [nseg]
  int *blkflag = nullptr;        // [nseg * nblk]
  int *segflag = nullptr;        // [nseg]
  Bins32 rws;                    // workspace: relay_nwb blocks of relay_wsz slots, or
  int64_t *rws_id = nullptr;     // relay_nws segments of cap slots
  int *rws_bs = nullptr;         // [relay_nws][2 (nx + 1)] a segment's new layout, count prefix
  int relay_wsz = 0, relay_nwb = 0, relay_nws = 0;
  double *fscratch = nullptr;    // 6 planes for the host-array boundary calls
  // restart files (paraio__init): the directory and the run's decomposition
  bool ioinit = false;
  std::string iodir;
  int io_nproc = 1, io_nrank = 0;
  // moment diagnostics: mom in the reference layout, per-tile node sums
  double *mom = nullptr, *mom_part = nullptr;
  long long cg_iters_total = 0;

  // implicit solve: plain CG (the reference's) or flexible CG with the
  // learned preconditioner, plain CG as fallback and judge
  int solver = WM_SOLVER_CG;
  int fcg_max = 50;              // flexible-CG iterations before the fallback
  // iterations queued before the first host check: the previous solve's
  // count, so a solve that converges as the last one did needs one check
  int fcg_hint = 2, cg_hint = 4;
  bool x0_fresh = false;         // cg_x0 holds df's solution planes (the warm start)
  wm_pc *pc = nullptr;
  double *cg_z = nullptr, *cg_q = nullptr, *cg_x0 = nullptr;   // 3 planes nx*ny each
  double *pc_sc = nullptr;       // s[3], 1/s[3] of the current residuals
  void *fcg_s = nullptr;         // flexible-CG scalars [3]
};

void wm_set_error(wm_ctx *ctx, int code, const char *fmt, ...);
int wm_alloc_device(wm_ctx *ctx);
int wm_check_derr(wm_ctx *ctx, const char *where);
int wm_fetch_derr(wm_ctx *ctx, int *out);     // device error bits, cleared
int wm_report_derr(wm_ctx *ctx, int h, const char *where);

// ---------------------------------------------------- kernel launch layer
// Particles (wm_particle.hip)
int wm_k_cell_fields(wm_ctx *ctx);
int wm_k_push(wm_ctx *ctx, int src, int dst);
int wm_k_deposit(wm_ctx *ctx, int oldb, int newb);
int wm_k_push_deposit(wm_ctx *ctx, int src, int dst);
int wm_k_deposit_gather(wm_ctx *ctx);
// Fields (wm_field.hip)
int wm_k_bc_curre(wm_ctx *ctx, double *uj);
XSide wm_xside(const wm_ctx *ctx);
int wm_k_bc_dfield(wm_ctx *ctx, double *df);
int wm_k_bc_phi(wm_ctx *ctx, double *phi, int l);   // phi(nxs-1:nxe+1, nys-1:nye+1)
int wm_k_field_solve(wm_ctx *ctx);   // RHS .. uf += df (after the current is ready)
int wm_k_energy(wm_ctx *ctx, int bank, double *out_host);
// Particle boundaries (wm_bc.hip)
int wm_k_particle_x(wm_ctx *ctx, int bank);
// xmode: 0 y only, 1 with the periodic x wrap fused, 2 with the walls' mirrors fused
int wm_k_particle_y(wm_ctx *ctx, int bank, int xmode);
// The resident step's particle_y and sort: pass 1, then pass 2 fused with
// the sort, from bank src into bank dst (the ID arrays follow); mid is
// recorded between the two
int wm_k_particle_y_sort(wm_ctx *ctx, int src, int dst, int xmode, hipEvent_t mid);
int wm_k_injection(wm_ctx *ctx, int bank, double u0);   // boundary_shock__injection
double wm_injection_xend(const wm_ctx *ctx, double u0);
// Sort (wm_sort.hip)
int wm_k_sort(wm_ctx *ctx, int src, int dst);
// FP32 form, cell bins (wm_sort.hip): dense rows <-> bins, movers into
// their destination bins, bin counts, row counts
int wm_k_stage_to_bins(wm_ctx *ctx, int seg0, int nseg, int w);   // staged AoS rows -> bins
int wm_k_bins_to_stage(wm_ctx *ctx, int seg0, int nseg, int w);   // bins -> staged AoS rows
int wm_k_bin_rows(wm_ctx *ctx);      // np2 and cumcnt of the bins
// The device-resident state row by row in the reference's record layout
// (wm_api.hip), for the restart files: batches of segments [s0, s1) of at
// most 2^26 doubles (wm_batch_end, w records per segment), out of the state
// into host h, or from host h into a state begun by wm_rows_begin and
// finished by wm_rows_done (FP64: the rows sorted by cell, as sort__bucket
// leaves them; FP32: the bins are sorted by construction).
int wm_batch_end(const wm_ctx *ctx, const int *cnt, int s0, int *w);
int wm_state_np2(wm_ctx *ctx, int *np2_host);
int wm_rows_out(wm_ctx *ctx, int s0, int s1, int w, double *h);
int wm_rows_begin(wm_ctx *ctx);
int wm_rows_in(wm_ctx *ctx, int s0, int s1, int w, const int *cnt_host, const double *h);
int wm_rows_done(wm_ctx *ctx, const int *np2_host);
int wm_field_out(wm_ctx *ctx, double *uf);
int wm_field_in(wm_ctx *ctx, const double *uf);
int wm_settle_movers(wm_ctx *ctx);   // the last push's movers into their bins, if pending
int wm_k_fill_check(wm_ctx *ctx);    // projected bin counts after a push; lists blocks to re-lay
int wm_k_id_checksum(wm_ctx *ctx, unsigned long long *out_dev);   // [sum, xor] of the IDs
int wm_relayout(wm_ctx *ctx);        // listed blocks (segments) laid out anew, on the device
// FP32 form (wm_particle32.hip)
int wm_k_cell_fields32(wm_ctx *ctx);
int wm_k_push_deposit32(wm_ctx *ctx);   // in place on the bins (pending movers first)
int wm_push32_pool();                     // movers one push workgroup lists
int wm_k_deposit_gather32(wm_ctx *ctx);
int wm_k_energy_kin32(wm_ctx *ctx, int bank, double *part, int nchunk);
// Moment diagnostics (wm_mom.hip): CIC moments of bank 0 with cumcnt (FP64)
// or of the settled bins (FP32) into ctx->mom (allocated by wm_mom_buffers on
// first use), u first advanced over half a step if accl; the half-step
// update alone into bank dst; the periodic fold
size_t wm_mom_elems(const wm_ctx *ctx);
int wm_mom_buffers(wm_ctx *ctx);
int wm_k_moments(wm_ctx *ctx, bool accl);
int wm_k_accl(wm_ctx *ctx, int src, int dst);
int wm_k_bc_mom(wm_ctx *ctx, double *mom);
// Learned preconditioner (wm_pc.hip)
int wm_pc_nparam();
int wm_pc_alloc(wm_ctx *ctx);
void wm_pc_free(wm_ctx *ctx);
int wm_pc_init_params(wm_ctx *ctx, uint64_t seed);
int wm_pc_set_params_dev(wm_ctx *ctx, const float *theta_host, int n);
int wm_pc_get_params_dev(wm_ctx *ctx, float *theta_host, int n);
int wm_pc_apply_host(wm_ctx *ctx, const double *r, double *z);
// z = M(r) on device images of at most WM_PC_NOALLOC_CELLS pixels, activations
// in a static device array (no allocation)
#define WM_PC_NOALLOC_CELLS (512 * 512)
int wm_pc_apply_noalloc_dev(wm_ctx *ctx, const double *r, double *z, int nx, int ny);
size_t wm_pc_noalloc_bytes();
int wm_pc_grad_host(wm_ctx *ctx, const double *r, const double *e, float *grad, double *loss);
int wm_pc_collect_cfg(wm_ctx *ctx, int nslot, int per_solve);
bool wm_pc_collecting(const wm_ctx *ctx);
int wm_pc_snapshot(wm_ctx *ctx, int k);
int wm_pc_snapshot_fin(wm_ctx *ctx);
int wm_pc_pairs(const wm_ctx *ctx);
int wm_pc_train_dev(wm_ctx *ctx, int steps, double lr, uint64_t seed, double *loss_out);
// z = M(r) inside flexible CG; also leaves per-tile partials of r.z and
// z.q (q ignored on a component's first iteration) in dots[l][2][tiles]
int wm_pc_apply_dev(wm_ctx *ctx, const double *r, const double *sc, double *z, const double *q,
                    double *dots, const FcgScal *S);
int wm_pc_dot_tiles(int nx, int ny);
bool wm_pc_ready(const wm_ctx *ctx);
void wm_pc_set_stats(wm_ctx *ctx, const int *fcg, const int *fb);
void wm_pc_get_stats(const wm_ctx *ctx, int *fcg, int *fb);
// Synthetic initial state (wm_init.hip)
int wm_k_init_weibel(wm_ctx *ctx, uint64_t seed, int ppc, double vth, double udrift);
int wm_k_init_weibel_bins(wm_ctx *ctx, uint64_t seed, int ppc, double vth, double udrift);

// --------------------------------------------------------- small helpers
static inline int wm_cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

#define WM_BLOCK 256
// phases timed by wm_step: cell-centred fields, push+deposit, current
// gather+fold, field solve, particle boundaries, sort
#define WM_NPHASE 6
