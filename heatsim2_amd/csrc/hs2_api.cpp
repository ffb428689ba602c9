// C ABI of heatsim2_amd (declared in include/heatsim2_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/heatsim2_amd.h"
#include "hs2_adjoint.h"
#include "hs2_batch.h"
#include "hs2_launch.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

int hs2::fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

namespace {
using hs2::fail;

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(HS2_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

enum { MODE_NONE = 0, MODE_CLASS = 1, MODE_DENSE = 2, MODE_ASYM = 3 };

constexpr int kCheckBlocks = 256;
constexpr int kMaxBlockThreads = 512;  // __launch_bounds__ of the strided kernel

}  // namespace

struct hs2_plan {
  hs2_grid_desc grid{};
  int n = 0;
  int device = 0;
  int mode = MODE_NONE;
  hs2::ClassView cv{};
  hs2::CombView comb{};         // x sweep's combined class plane (class mode)
  bool have_comb = false;       // built (the class triples number <= 256)
  bool use_comb = true;         // hs2_plan_set_option(HS2_OPT_COMBINED_X)
  uint8_t *d_comb = nullptr;    // n combined ids
  double *d_ctab = nullptr;     // 12 doubles per id
  hs2::DenseView dv{};
  hs2::AsymView av{};
  int ncls[3] = {0, 0, 0};
  uint8_t *d_cls = nullptr;     // 3n class ids (class mode)
  double *d_tab = nullptr;      // class tables
  uint8_t *d_lut = nullptr;     // classification lookup tables
  double *d_dense = nullptr;    // plan-built dense arrays: M, gz, gy, gx (4n)
  uint8_t *d_densew = nullptr;  // plan-built dense row weights (n)
  double *d_ws = nullptr;       // n doubles: T* then T**
  double *d_vtab = nullptr;     // kMaxVtab doubles (hs2_step)
  double *d_vsteps = nullptr;   // per-step source tables (hs2_run)
  size_t vsteps_cap = 0;        // doubles
  int *d_check = nullptr;       // per-block maxima of the maps (validation)
  hs2::XConfig xcfg{0, 0, 0};
  hs2::SConfig scfg[2] = {{0, 0}, {0, 0}};  // [0] z, [1] y
};

enum { KIND_X2 = 0, KIND_Y = 1, KIND_Z = 2, KIND_AX = 3, KIND_CELL = 4, NKIND = 5 };

struct hs2_batch {
  int ninst = 0, K = 0, maxsteps = 0, device = 0;
  std::vector<hs2::BatchInst> inst;
  long long N = 0, H = 0, CK = 0;  // cells, history values, checkpoint values
  hs2::BatchInst *d_inst = nullptr;
  int *d_ctr = nullptr;
  int nblk[NKIND] = {};
  int *d_blk_inst[NKIND] = {};
  int *d_blk0[NKIND] = {};
  hs2::DenseView dv{};
  bool have_coef = false;
  int use_graph = 1, chunk = 8;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t fwd_exec = nullptr, bwd_exec = nullptr;
  std::vector<const void *> fwd_key, bwd_key;
};

namespace {

}  // namespace

// Untuned y / z shape: segments of up to 16 cells (32-cell segments spill
// registers), and blocks of about 256 threads, as the tuner picks at 256 and
// 512 cells per line.
int hs2::default_strided(int L, hs2::SConfig *cfg) {
  int s;
  if (L <= 4) s = 2;
  else if (L <= 16) s = 4;
  else if (L <= 64) s = 8;
  else if (L <= 1024) s = 16;
  else return 0;
  const int P = (L + s - 1) / s;
  int w = 64;
  while (w > 8 && P * w > 256) w >>= 1;
  if (P * w > kMaxBlockThreads) return 0;
  *cfg = {s, w};
  return 1;
}

namespace {
using hs2::default_strided;

bool strided_ok(int L, hs2::SConfig c) {
  const bool seg_ok = c.seg == 2 || c.seg == 4 || c.seg == 8 || c.seg == 16 || c.seg == 32;
  const bool cw_ok = c.cw == 8 || c.cw == 16 || c.cw == 32 || c.cw == 64;
  if (!seg_ok || !cw_ok) return false;
  const int P = (L + c.seg - 1) / c.seg;
  return P * c.cw <= kMaxBlockThreads;
}

bool x_ok(int nx, hs2::XConfig c) {
  if (c.ty != 0 && (c.lpb == 0 || c.ty < 0 || c.ty >= c.lpb || c.lpb % c.ty != 0)) return false;
  static const int combos[][3] = {
      {1, 16, 0}, {2, 16, 0}, {4, 16, 0}, {8, 16, 0}, {16, 16, 0}, {2, 32, 0}, {4, 32, 0},
      {8, 32, 0}, {16, 32, 0}, {2, 64, 0}, {4, 64, 0}, {8, 64, 0}, {16, 64, 0},
      {1, 16, 8}, {2, 16, 8}, {4, 16, 8}, {8, 16, 8}, {16, 16, 8}, {4, 16, 16}, {8, 16, 16},
      {16, 16, 4}, {16, 16, 16}, {8, 32, 4}, {16, 32, 4}, {16, 32, 8}, {16, 32, 2}, {16, 64, 2},
      {8, 64, 2}, {16, 64, 4}};
  for (auto &cb : combos)
    if (cb[0] == c.cpl && cb[1] == c.g && cb[2] == c.lpb) return c.cpl * c.g >= nx;
  return false;
}

// per-block maximum of a uint8 map (out[blockIdx.x]); blockDim.x == 256
__global__ void k_map_max(const uint8_t *a, size_t n, int *out) {
  __shared__ int red[256];
  int m = 0;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    m = max(m, (int)a[t]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// dense arrays from the maps and the full conductance table (used when an
// axis has more than 256 distinct coefficient classes)
__global__ void k_dense_from_maps(hs2::Grid g, const uint8_t *__restrict__ mat,
                                  const uint8_t *__restrict__ bz, const uint8_t *__restrict__ by,
                                  const uint8_t *__restrict__ bx, const double *__restrict__ gtab,
                                  const double *__restrict__ mtab, const uint8_t *__restrict__ comp,
                                  int nmat, double *M, uint8_t *w, double *gz, double *gy,
                                  double *gx) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int n = nx * ny * nz, sxy = nx * ny;
  const size_t ax = (size_t)nmat * nmat;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int k = c / sxy, rem = c - k * sxy, j = rem / nx, i = rem - j * nx;
    const int m = mat[c];
    M[c] = mtab[m];
    w[c] = comp[m];
    gz[c] = k > 0 ? gtab[(bz[c] * 3 + 0) * ax + (size_t)mat[c - sxy] * nmat + m] : 0.0;
    gy[c] = j > 0 ? gtab[(by[(k * (ny + 1) + j) * nx + i] * 3 + 1) * ax +
                         (size_t)mat[c - nx] * nmat + m]
                  : 0.0;
    gx[c] = i > 0 ? gtab[(bx[(k * ny + j) * (nx + 1) + i] * 3 + 2) * ax +
                         (size_t)mat[c - 1] * nmat + m]
                  : 0.0;
  }
}

int map_max(const uint8_t *d_map, size_t n, int *d_scratch, int *result) {
  hipLaunchKernelGGL(k_map_max, dim3(kCheckBlocks), dim3(256), 0, 0, d_map, n, d_scratch);
  HIP_TRY(hipGetLastError());
  int part[kCheckBlocks];
  HIP_TRY(hipMemcpy(part, d_scratch, sizeof part, hipMemcpyDeviceToHost));
  *result = *std::max_element(part, part + kCheckBlocks);
  return HS2_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

hs2::Grid grid_of(const hs2_plan *p) { return hs2::Grid{p->grid.nz, p->grid.ny, p->grid.nx}; }

hipError_t sweep_x(const hs2_plan *p, const double *Tn, double *Ts, const uint8_t *vmap,
                   const double *vtab, const double *sfield, hipStream_t s, bool ry = false) {
  const hs2::Grid g = grid_of(p);
  if (p->mode == MODE_CLASS && p->have_comb && p->use_comb && !vmap && !sfield &&
      hs2::x_ry_ok(p->xcfg, g))
    return hs2::launch_sweep_x(p->xcfg, g, p->comb, Tn, Ts, s, ry);
  if (p->mode == MODE_CLASS)
    return hs2::launch_sweep_x(p->xcfg, g, p->cv, Tn, Ts, vmap, vtab, sfield, s, ry);
  if (p->mode == MODE_ASYM)
    return hs2::launch_sweep_x(p->xcfg, g, p->av, Tn, Ts, vmap, vtab, sfield, s);
  return hs2::launch_sweep_x(p->xcfg, g, p->dv, Tn, Ts, vmap, vtab, sfield, s);
}

hipError_t sweep_s(const hs2_plan *p, int axis, const double *Tn, const double *Y, double *X,
                   double *surf, hipStream_t s, bool ry = false) {
  const hs2::SConfig c = p->scfg[axis];
  const hs2::Grid g = grid_of(p);
  const double dz = p->grid.dz;
  if (p->mode == MODE_CLASS)
    return hs2::launch_sweep_strided(c, g, p->cv, axis, Tn, Y, X, surf, dz, s, ry);
  if (p->mode == MODE_ASYM)
    return hs2::launch_sweep_strided(c, g, p->av, axis, Tn, Y, X, surf, dz, s);
  return hs2::launch_sweep_strided(c, g, p->dv, axis, Tn, Y, X, surf, dz, s);
}

// the x sweep hands the y sweep its right-hand side instead of T* (one field
// read less) where the plan and the step allow it
bool use_ry(const hs2_plan *p, bool src) {
  return p->mode == MODE_CLASS && !src && hs2::x_ry_ok(p->xcfg, grid_of(p));
}

int run_step(hs2_plan *p, const double *Tin, double *Tout, const uint8_t *vmap,
             const double *d_vtab, const double *sfield, double *surf, hipStream_t s) {
  const bool ry = use_ry(p, vmap || sfield);
  HIP_TRY(sweep_x(p, Tin, p->d_ws, vmap, d_vtab, sfield, s, ry));
  HIP_TRY(sweep_s(p, 1, Tin, p->d_ws, p->d_ws, nullptr, s, ry));
  HIP_TRY(sweep_s(p, 0, Tin, p->d_ws, Tout, surf, s));
  return HS2_OK;
}

// one step writing its intermediate states: T* -> Tx, T** -> Ty, T^{n+1} -> Tout
int step_states(hs2_plan *p, const double *Tin, double *Tx, double *Ty, double *Tout,
                const double *sfield, double *surf, hipStream_t s) {
  HIP_TRY(sweep_x(p, Tin, Tx, nullptr, nullptr, sfield, s));
  HIP_TRY(sweep_s(p, 1, Tin, Tx, Ty, nullptr, s));
  HIP_TRY(sweep_s(p, 0, Tin, Ty, Tout, surf, s));
  return HS2_OK;
}

hipError_t adj_line(const hs2_plan *p, int axis, const double *R, int scale_m, double *U,
                    hipStream_t s) {
  return p->mode == MODE_CLASS ? hs2::launch_adj_line(grid_of(p), p->cv, axis, R, scale_m, U, s)
                               : hs2::launch_adj_line(grid_of(p), p->dv, axis, R, scale_m, U, s);
}

hipError_t adj_stencil(const hs2_plan *p, const hs2::AdjFields &f, hipStream_t s) {
  return p->mode == MODE_CLASS ? hs2::launch_adj_stencil(grid_of(p), p->cv, f, s)
                               : hs2::launch_adj_stencil(grid_of(p), p->dv, f, s);
}

// step -> source index (-1: none); validates the schedule
int source_schedule(int nsteps, const int32_t *h_src_step, int nsrc, const double *d_src,
                    std::vector<int> *out, const char *who) {
  out->assign((size_t)nsteps, -1);
  if (nsrc < 0 || (nsrc > 0 && (!h_src_step || !d_src)))
    return fail(HS2_E_ARG, "%s: bad source arguments", who);
  if (nsrc > 0 && !aligned16(d_src)) return fail(HS2_E_ARG, "%s: d_src must be 16-byte aligned", who);
  for (int s = 0; s < nsrc; s++) {
    const int st = h_src_step[s];
    if (st < 0 || st >= nsteps) return fail(HS2_E_ARG, "%s: source %d at step %d outside [0,%d)", who, s, st, nsteps);
    if ((*out)[st] >= 0) return fail(HS2_E_ARG, "%s: two sources at step %d", who, st);
    (*out)[st] = s;
  }
  return HS2_OK;
}

void free_coefficients(hs2_plan *p) {
  (void)hipFree(p->d_cls);
  (void)hipFree(p->d_tab);
  (void)hipFree(p->d_lut);
  (void)hipFree(p->d_dense);
  (void)hipFree(p->d_densew);
  p->d_cls = nullptr;
  p->d_tab = nullptr;
  p->d_lut = nullptr;
  p->d_dense = nullptr;
  p->d_densew = nullptr;
  (void)hipFree(p->d_comb);
  (void)hipFree(p->d_ctab);
  p->d_comb = nullptr;
  p->d_ctab = nullptr;
  p->have_comb = false;
  p->mode = MODE_NONE;
}

}  // namespace

// ------------------------------------------------------------- batches
namespace {

hs2::BatchMap batch_map(const hs2_batch *b, int kind, int rule, int j) {
  return hs2::BatchMap{b->d_inst, b->d_blk_inst[kind], b->d_blk0[kind], b->d_ctr, rule, j, b->K};
}

int blocks_of(const hs2::BatchInst &I, int kind) {
  switch (kind) {
    case KIND_X2: return (I.nz * I.ny + hs2::kBatchXLines - 1) / hs2::kBatchXLines;
    case KIND_Y: return (I.nx * I.nz + I.cw[1] - 1) / I.cw[1];
    case KIND_Z: return (I.nx * I.ny + I.cw[0] - 1) / I.cw[0];
    case KIND_AX: return (I.nz * I.ny + I.cw[2] - 1) / I.cw[2];
    default: return (int)(((long long)I.nz * I.ny * I.nx + 255) / 256);
  }
}

void batch_free(hs2_batch *b) {
  if (b->fwd_exec) (void)hipGraphExecDestroy(b->fwd_exec);
  if (b->bwd_exec) (void)hipGraphExecDestroy(b->bwd_exec);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  (void)hipFree(b->d_inst);
  (void)hipFree(b->d_ctr);
  for (int k = 0; k < NKIND; k++) {
    (void)hipFree(b->d_blk_inst[k]);
    (void)hipFree(b->d_blk0[k]);
  }
}

// work layout: ws (N) | checkpoints (CK) | segment T^t (K+1 N) | T* (K N) | T** (K N) |
// ux uy uz lam (4 N)
struct BatchWork {
  double *ws, *ck, *seg, *segx, *segy, *ux, *uy, *uz, *lam;
};
long long batch_work_doubles(const hs2_batch *b) { return b->N * (3LL * b->K + 6) + b->CK; }
BatchWork batch_work(const hs2_batch *b, double *w) {
  BatchWork r;
  r.ws = w;
  r.ck = r.ws + b->N;
  r.seg = r.ck + b->CK;
  r.segx = r.seg + (size_t)(b->K + 1) * b->N;
  r.segy = r.segx + (size_t)b->K * b->N;
  r.ux = r.segy + (size_t)b->K * b->N;
  r.uy = r.ux + b->N;
  r.uz = r.uy + b->N;
  r.lam = r.uz + b->N;
  return r;
}

// forward: `steps` steps (ckpt store, x, y, z with the readout, tick)
hipError_t emit_forward(hs2_batch *b, double *T, const double *src, double *hist,
                        const BatchWork &w, int steps, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (int c = 0; c < steps && e == hipSuccess; c++) {
    e = hs2::launch_batch_x2(batch_map(b, KIND_X2, hs2::RULE_FWD, 0), b->nblk[KIND_X2], b->dv, T,
                             w.ws, src, w.ck, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_strided(batch_map(b, KIND_Y, hs2::RULE_FWD, 0), b->nblk[KIND_Y], 1,
                                    b->dv, T, w.ws, w.ws, nullptr, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_strided(batch_map(b, KIND_Z, hs2::RULE_FWD, 0), b->nblk[KIND_Z], 0,
                                    b->dv, T, w.ws, T, hist, s);
    if (e == hipSuccess) e = hs2::launch_batch_counter(b->d_ctr, 0, 1, 1, s);
  }
  return e;
}

// one backward segment: load checkpoints, recompute its K steps keeping every
// state, K adjoint steps, tick
hipError_t emit_backward_segment(hs2_batch *b, const double *src, const double *ghist,
                                 double *gM, double *gz, double *gy, double *gx, double *gsrc,
                                 const BatchWork &w, hipStream_t s) {
  const int K = b->K;
  const size_t N = (size_t)b->N;
  hipError_t e = hs2::launch_batch_ckpt_load(batch_map(b, KIND_CELL, hs2::RULE_SEG, 0),
                                             b->nblk[KIND_CELL], w.ck, w.seg, b->N, s);
  for (int j = 0; j < K && e == hipSuccess; j++) {
    const double *in = w.seg + (size_t)j * N;
    double *tx = w.segx + (size_t)j * N, *ty = w.segy + (size_t)j * N;
    e = hs2::launch_batch_x2(batch_map(b, KIND_X2, hs2::RULE_RECOMP, j), b->nblk[KIND_X2], b->dv,
                             in, tx, src, nullptr, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_strided(batch_map(b, KIND_Y, hs2::RULE_RECOMP, j), b->nblk[KIND_Y], 1,
                                    b->dv, in, tx, ty, nullptr, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_strided(batch_map(b, KIND_Z, hs2::RULE_RECOMP, j), b->nblk[KIND_Z], 0,
                                    b->dv, in, ty, w.seg + (size_t)(j + 1) * N, nullptr, s);
  }
  // adjoint steps, latest first: z, y, x transposed line solves from
  // lam = dJ/dT^{t+1}, then the stencil, which leaves dJ/dT^t in lam
  for (int j = 0; j < K && e == hipSuccess; j++) {
    const size_t q = (size_t)(K - 1 - j);
    e = hs2::launch_batch_adj_line(batch_map(b, KIND_Z, hs2::RULE_BWD, j), b->nblk[KIND_Z], 0,
                                   b->dv, w.lam, 0, w.uz, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_adj_line(batch_map(b, KIND_Y, hs2::RULE_BWD, j), b->nblk[KIND_Y], 1,
                                     b->dv, w.uz, 1, w.uy, s);
    if (e == hipSuccess)
      e = hs2::launch_batch_adj_line(batch_map(b, KIND_AX, hs2::RULE_BWD, j), b->nblk[KIND_AX], 2,
                                     b->dv, w.uy, 1, w.ux, s);
    if (e == hipSuccess) {
      hs2::AdjFields f{};
      f.Tn = w.seg + q * N; f.Tx = w.segx + q * N; f.Ty = w.segy + q * N;
      f.Tn1 = w.seg + (q + 1) * N;
      f.ux = w.ux; f.uy = w.uy; f.uz = w.uz;
      f.ghist = ghist;
      f.lam_out = w.lam;
      f.gM = gM;
      f.gg[0] = gz; f.gg[1] = gy; f.gg[2] = gx;
      f.gsrc = gsrc;
      e = hs2::launch_batch_adj_stencil(batch_map(b, KIND_CELL, hs2::RULE_BWD, j),
                                        b->nblk[KIND_CELL], b->dv, f, s);
    }
  }
  if (e == hipSuccess) e = hs2::launch_batch_counter(b->d_ctr, 1, 1, 1, s);
  return e;
}

// run `emit` `reps` times on b->stream, through a graph captured once per key
template <class Emit>
int batch_replay(hs2_batch *b, hipGraphExec_t *exec, std::vector<const void *> *key,
                 const std::vector<const void *> &want, int reps, Emit emit, const char *who) {
  hipStream_t s = b->stream;
  if (!b->use_graph) {
    for (int r = 0; r < reps; r++) HIP_TRY(emit(s));
    return HS2_OK;
  }
  if (!*exec || *key != want) {
    if (*exec) {
      HIP_TRY(hipGraphExecDestroy(*exec));
      *exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const hipError_t ce = emit(s);
    const hipError_t ee = hipStreamEndCapture(s, &graph);
    if (ce != hipSuccess || ee != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(HS2_E_HIP, "%s: graph capture failed: %s", who,
                  hipGetErrorString(ce != hipSuccess ? ce : ee));
    }
    const hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      *exec = nullptr;
      return fail(HS2_E_HIP, "%s: hipGraphInstantiate: %s", who, hipGetErrorString(ie));
    }
    *key = want;
  }
  for (int r = 0; r < reps; r++) HIP_TRY(hipGraphLaunch(*exec, s));
  return HS2_OK;
}

// The x sweep's combined class plane: one id per cell for its (z, y, x)
// class triple, when the triples number 256 or fewer (else the x sweep
// reads the three class planes).  tabs/off: the per-axis class tables.
int build_combined(hs2_plan *p, const std::vector<double> &tabs, const int off[4]) {
  constexpr size_t kWords = (size_t(1) << 24) / 32;
  const int n = p->n;
  const uint8_t *cz = p->cv.cls[0], *cy = p->cv.cls[1], *cx = p->cv.cls[2];
  uint32_t *d_bits = nullptr;
  int *d_rank = nullptr;
  std::vector<uint32_t> bits(kWords);
  std::vector<int> rank0(kWords);
  std::vector<uint32_t> keys;
  auto done = [&](int rc) {
    (void)hipFree(d_bits);
    (void)hipFree(d_rank);
    return rc;
  };
  hipError_t e = hipMalloc(&d_bits, kWords * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(d_bits, 0, kWords * sizeof(uint32_t));
  if (e == hipSuccess) e = hs2::launch_comb_mark(n, cz, cy, cx, d_bits, 0);
  if (e == hipSuccess)
    e = hipMemcpy(bits.data(), d_bits, kWords * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return done(fail(HS2_E_HIP, "combined class plane: %s", hipGetErrorString(e)));
  int cnt = 0;
  for (size_t w = 0; w < kWords; w++) {
    rank0[w] = cnt;
    for (uint32_t b = bits[w]; b; b &= b - 1, cnt++)
      keys.push_back((uint32_t)(w * 32 + __builtin_ctz(b)));
  }
  if (cnt > hs2::kMaxClass) return done(HS2_OK);  // the three planes serve
  std::vector<double> ctab((size_t)hs2::CombView::kStride * cnt, 0.0);
  for (int i = 0; i < cnt; i++) {
    const uint32_t id[3] = {keys[i] & 0xffu, (keys[i] >> 8) & 0xffu, keys[i] >> 16};  // z, y, x
    for (int a = 0; a < 3; a++) {
      const double *src = &tabs[off[a] + (size_t)id[a] * hs2::kTabStride];
      double *dst = &ctab[(size_t)i * hs2::CombView::kStride + 4 * (2 - a)];
      for (int k = 0; k < 3; k++) dst[k] = src[k];
    }
  }
  e = hipMalloc(&d_rank, kWords * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(d_rank, rank0.data(), kWords * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&p->d_comb, (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&p->d_ctab, sizeof(double) * ctab.size());
  if (e == hipSuccess)
    e = hipMemcpy(p->d_ctab, ctab.data(), sizeof(double) * ctab.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hs2::launch_comb_assign(n, cz, cy, cx, d_bits, d_rank, p->d_comb, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(p->d_comb);
    (void)hipFree(p->d_ctab);
    p->d_comb = nullptr;
    p->d_ctab = nullptr;
    return done(fail(HS2_E_HIP, "combined class plane: %s", hipGetErrorString(e)));
  }
  p->comb.comb = p->d_comb;
  p->comb.tab = p->d_ctab;
  p->comb.n = cnt;
  p->have_comb = true;
  return done(HS2_OK);
}

}  // namespace

extern "C" {

const char *hs2_last_error(void) { return g_last_error.c_str(); }

const char *hs2_version(void) { return "heatsim2_amd 0.5 (gfx950, fp64)"; }

int hs2_plan_create(const hs2_grid_desc *grid, hs2_plan **out) {
  if (!grid || !out) return fail(HS2_E_ARG, "hs2_plan_create: null argument");
  *out = nullptr;
  const hs2_grid_desc &G = *grid;
  if (G.nz < 1 || G.ny < 1 || G.nx < 1)
    return fail(HS2_E_ARG, "hs2_plan_create: grid %dx%dx%d must be positive", G.nz, G.ny, G.nx);
  if (!(G.dz > 0 && G.dy > 0 && G.dx > 0 && G.dt > 0))
    return fail(HS2_E_ARG, "hs2_plan_create: spacings and dt must be positive");
  const long long n = (long long)G.nz * G.ny * G.nx;
  if ((long long)(G.nz + 1) * G.ny * G.nx >= (1LL << 31) ||
      (long long)G.nz * G.ny * (G.nx + 1) >= (1LL << 31))
    return fail(HS2_E_UNSUPPORTED, "hs2_plan_create: grid of %lld cells exceeds 2^31 indices", n);
  const hs2::XConfig xc = hs2::x_config_for(G.nx, G.nz);
  if (!xc.cpl) return fail(HS2_E_UNSUPPORTED, "hs2_plan_create: nx=%d > 1024 unsupported", G.nx);
  hs2_plan *p = new hs2_plan();
  p->grid = G;
  p->n = (int)n;
  p->xcfg = xc;
  if (!default_strided(G.nz, &p->scfg[0]) || !default_strided(G.ny, &p->scfg[1])) {
    delete p;
    return fail(HS2_E_UNSUPPORTED, "hs2_plan_create: ny=%d / nz=%d > 1024 unsupported", G.ny, G.nz);
  }
  hipError_t e = hipGetDevice(&p->device);
  if (e == hipSuccess) e = hipMalloc(&p->d_ws, sizeof(double) * (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&p->d_vtab, sizeof(double) * hs2::kMaxVtab);
  if (e == hipSuccess) e = hipMalloc(&p->d_check, sizeof(int) * kCheckBlocks);
  if (e != hipSuccess) {
    hs2_plan_destroy(p);
    return fail(HS2_E_ALLOC, "hs2_plan_create: %s", hipGetErrorString(e));
  }
  *out = p;
  return HS2_OK;
}

void hs2_plan_destroy(hs2_plan *p) {
  if (!p) return;
  free_coefficients(p);
  (void)hipFree(p->d_ws);
  (void)hipFree(p->d_vtab);
  (void)hipFree(p->d_vsteps);
  (void)hipFree(p->d_check);
  delete p;
}

int hs2_plan_set_compact(hs2_plan *p, const uint8_t *d_mat, const uint8_t *d_bz,
                         const uint8_t *d_by, const uint8_t *d_bx, int nmat, int nb,
                         const double *h_gtab, const double *h_mtab, const uint8_t *h_compute) {
  if (!p || !d_mat || !d_bz || !d_by || !d_bx || !h_gtab || !h_mtab || !h_compute)
    return fail(HS2_E_ARG, "hs2_plan_set_compact: null argument");
  if (nmat < 1 || nmat > 256 || nb < 1 || nb > 256)
    return fail(HS2_E_ARG, "hs2_plan_set_compact: nmat=%d nb=%d must be in 1..256", nmat, nb);
  const size_t lutlen = (size_t)nb * nmat * nmat;
  const size_t glen = 3 * lutlen;
  for (int m = 0; m < nmat; m++) {
    if (h_compute[m] > 1) return fail(HS2_E_ARG, "hs2_plan_set_compact: compute flag not 0/1");
    if (!(h_mtab[m] > 0.0) || h_mtab[m] > 1e300)
      return fail(HS2_E_ARG, "hs2_plan_set_compact: M[%d]=%g must be positive", m, h_mtab[m]);
  }
  for (size_t t = 0; t < glen; t++)
    if (!(h_gtab[t] >= 0.0) || h_gtab[t] > 1e300)
      return fail(HS2_E_ARG, "hs2_plan_set_compact: conductance table entry %zu = %g invalid", t,
                  h_gtab[t]);
  const hs2_grid_desc &G = p->grid;
  const size_t nzf = (size_t)(G.nz + 1) * G.ny * G.nx, nyf = (size_t)G.nz * (G.ny + 1) * G.nx,
               nxf = (size_t)G.nz * G.ny * (G.nx + 1);
  int mx[4], rc;
  if ((rc = map_max(d_mat, (size_t)p->n, p->d_check, &mx[0])) ||
      (rc = map_max(d_bz, nzf, p->d_check, &mx[1])) ||
      (rc = map_max(d_by, nyf, p->d_check, &mx[2])) ||
      (rc = map_max(d_bx, nxf, p->d_check, &mx[3])))
    return rc;
  if (mx[0] >= nmat)
    return fail(HS2_E_ARG, "hs2_plan_set_compact: material map holds %d >= nmat=%d", mx[0], nmat);
  for (int t = 1; t < 4; t++)
    if (mx[t] >= nb)
      return fail(HS2_E_ARG, "hs2_plan_set_compact: %s face map holds %d >= nb=%d",
                  t == 1 ? "z" : (t == 2 ? "y" : "x"), mx[t], nb);

  // Coefficient classes per axis: a cell's entry is (M, w, g of its minus
  // face); g is 0 on the domain edge.  Distinct entries are numbered.
  std::vector<uint8_t> lut(3 * lutlen + 3 * (size_t)nmat, 0);
  std::vector<double> tabs;
  int off[4] = {0, 0, 0, 0};
  bool fits = true;
  for (int a = 0; a < 3 && fits; a++) {
    std::map<std::tuple<double, double, double>, int> ids;
    std::vector<std::tuple<double, double, double>> order;
    auto id_of = [&](double M, double w, double gv) -> int {
      auto key = std::make_tuple(M, w, gv);
      auto it = ids.find(key);
      if (it != ids.end()) return it->second;
      const int id = (int)order.size();
      ids.emplace(key, id);
      order.push_back(key);
      return id;
    };
    for (int m = 0; m < nmat; m++) {
      const int id = id_of(h_mtab[m], h_compute[m] ? 1.0 : 0.0, 0.0);
      if (id >= hs2::kMaxClass) { fits = false; break; }
      lut[3 * lutlen + (size_t)a * nmat + m] = (uint8_t)id;
    }
    for (int b = 0; b < nb && fits; b++)
      for (int ma = 0; ma < nmat && fits; ma++)
        for (int m = 0; m < nmat; m++) {
          const double gv = h_gtab[(((size_t)b * 3 + a) * nmat + ma) * nmat + m];
          const int id = id_of(h_mtab[m], h_compute[m] ? 1.0 : 0.0, gv);
          if (id >= hs2::kMaxClass) { fits = false; break; }
          lut[(size_t)a * lutlen + ((size_t)b * nmat + ma) * nmat + m] = (uint8_t)id;
        }
    off[a] = (int)tabs.size();
    for (auto &e : order) {
      tabs.push_back(std::get<0>(e));
      tabs.push_back(std::get<1>(e));
      tabs.push_back(std::get<2>(e));
      tabs.push_back(0.0);
    }
    p->ncls[a] = (int)order.size();
  }
  off[3] = (int)tabs.size();
  free_coefficients(p);
  const hs2::Grid g = grid_of(p);
  const size_t n = (size_t)p->n;
  if (fits) {
    HIP_TRY(hipMalloc(&p->d_cls, 3 * n));
    HIP_TRY(hipMalloc(&p->d_tab, sizeof(double) * tabs.size()));
    HIP_TRY(hipMalloc(&p->d_lut, lut.size()));
    HIP_TRY(hipMemcpy(p->d_tab, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->d_lut, lut.data(), lut.size(), hipMemcpyHostToDevice));
    uint8_t *cz = p->d_cls, *cy = p->d_cls + n, *cx = p->d_cls + 2 * n;
    HIP_TRY(hs2::launch_classify(g, d_mat, d_bz, d_by, d_bx, p->d_lut, p->d_lut + 3 * lutlen, nmat,
                                 (int)lutlen, cz, cy, cx, 0));
    HIP_TRY(hipDeviceSynchronize());
    p->cv.cls[0] = cz;
    p->cv.cls[1] = cy;
    p->cv.cls[2] = cx;
    p->cv.tab = p->d_tab;
    for (int t = 0; t < 4; t++) p->cv.off[t] = off[t];
    p->mode = MODE_CLASS;
    return build_combined(p, tabs, off);
  }
  // more than 256 classes on an axis: per-cell float64 coefficients
  double *d_gtab = nullptr, *d_mtab = nullptr;
  uint8_t *d_comp = nullptr;
  hipError_t e = hipMalloc(&p->d_dense, sizeof(double) * 4 * n);
  if (e == hipSuccess) e = hipMalloc(&p->d_densew, n);
  if (e == hipSuccess) e = hipMalloc(&d_gtab, sizeof(double) * glen);
  if (e == hipSuccess) e = hipMalloc(&d_mtab, sizeof(double) * nmat);
  if (e == hipSuccess) e = hipMalloc(&d_comp, nmat);
  if (e == hipSuccess) e = hipMemcpy(d_gtab, h_gtab, sizeof(double) * glen, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_mtab, h_mtab, sizeof(double) * nmat, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_comp, h_compute, nmat, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_dense_from_maps, dim3(blocks), dim3(256), 0, 0, g, d_mat, d_bz, d_by, d_bx,
                       d_gtab, d_mtab, d_comp, nmat, p->d_dense, p->d_densew, p->d_dense + n,
                       p->d_dense + 2 * n, p->d_dense + 3 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d_gtab);
  (void)hipFree(d_mtab);
  (void)hipFree(d_comp);
  if (e != hipSuccess) {
    free_coefficients(p);
    return fail(HS2_E_HIP, "hs2_plan_set_compact (dense fallback): %s", hipGetErrorString(e));
  }
  p->dv.M = p->d_dense;
  p->dv.w = p->d_densew;
  p->dv.g[0] = p->d_dense + n;
  p->dv.g[1] = p->d_dense + 2 * n;
  p->dv.g[2] = p->d_dense + 3 * n;
  p->mode = MODE_DENSE;
  return HS2_OK;
}

int hs2_plan_set_dense_asym(hs2_plan *p, const double *d_M, const uint8_t *d_w,
                            const double *d_gzm, const double *d_gym, const double *d_gxm,
                            const double *d_gzp, const double *d_gyp, const double *d_gxp) {
  const double *g[6] = {d_gzm, d_gym, d_gxm, d_gzp, d_gyp, d_gxp};
  if (!p || !d_M || !d_w) return fail(HS2_E_ARG, "hs2_plan_set_dense_asym: null argument");
  for (const double *a : g)
    if (!a) return fail(HS2_E_ARG, "hs2_plan_set_dense_asym: null argument");
  if (!aligned16(d_M) || !aligned16(d_w))
    return fail(HS2_E_ARG, "hs2_plan_set_dense_asym: arrays must be 16-byte aligned");
  for (const double *a : g)
    if (!aligned16(a)) return fail(HS2_E_ARG, "hs2_plan_set_dense_asym: arrays must be 16-byte aligned");
  free_coefficients(p);
  p->av.M = d_M;
  p->av.w = d_w;
  for (int a = 0; a < 3; a++) {
    p->av.g[a] = g[a];
    p->av.gp[a] = g[3 + a];
  }
  p->mode = MODE_ASYM;
  return HS2_OK;
}

int hs2_plan_set_dense(hs2_plan *p, const double *d_M, const uint8_t *d_w, const double *d_gz,
                       const double *d_gy, const double *d_gx) {
  if (!p || !d_M || !d_w || !d_gz || !d_gy || !d_gx)
    return fail(HS2_E_ARG, "hs2_plan_set_dense: null argument");
  if (!aligned16(d_M) || !aligned16(d_w) || !aligned16(d_gz) || !aligned16(d_gy) ||
      !aligned16(d_gx))
    return fail(HS2_E_ARG, "hs2_plan_set_dense: arrays must be 16-byte aligned");
  free_coefficients(p);
  p->dv.M = d_M;
  p->dv.w = d_w;
  p->dv.g[0] = d_gz;
  p->dv.g[1] = d_gy;
  p->dv.g[2] = d_gx;
  p->mode = MODE_DENSE;
  return HS2_OK;
}

int hs2_plan_info(const hs2_plan *p, int *info) {
  if (!p || !info) return fail(HS2_E_ARG, "hs2_plan_info: null argument");
  info[10] = p->xcfg.lpb;
  info[0] = p->mode;
  info[1] = p->ncls[0];
  info[2] = p->ncls[1];
  info[3] = p->ncls[2];
  info[4] = p->xcfg.cpl;
  info[5] = p->xcfg.g;
  info[6] = p->scfg[1].seg;
  info[7] = p->scfg[1].cw;
  info[8] = p->scfg[0].seg;
  info[9] = p->scfg[0].cw;
  info[11] = p->xcfg.ty;
  info[12] = use_ry(p, false) ? 1 : 0;
  info[13] = p->mode == MODE_CLASS && p->have_comb && p->use_comb &&
             hs2::x_ry_ok(p->xcfg, grid_of(p)) ? p->comb.n : 0;
  return HS2_OK;
}

int hs2_plan_set_config(hs2_plan *p, int x_cpl, int x_lanes, int x_lines, int x_lines_y,
                        int y_seg, int y_lines, int z_seg, int z_lines) {
  if (!p) return fail(HS2_E_ARG, "hs2_plan_set_config: null plan");
  const hs2::XConfig xc{x_cpl, x_lanes, x_lines, x_lines_y};
  const hs2::SConfig yc{y_seg, y_lines}, zc{z_seg, z_lines};
  if (!x_ok(p->grid.nx, xc))
    return fail(HS2_E_ARG, "hs2_plan_set_config: x config %dx%dx%d (%d rows) invalid for nx=%d",
                x_cpl, x_lanes, x_lines, x_lines_y, p->grid.nx);
  if (!strided_ok(p->grid.ny, yc))
    return fail(HS2_E_ARG, "hs2_plan_set_config: y config seg=%d lines=%d invalid for ny=%d",
                y_seg, y_lines, p->grid.ny);
  if (!strided_ok(p->grid.nz, zc))
    return fail(HS2_E_ARG, "hs2_plan_set_config: z config seg=%d lines=%d invalid for nz=%d",
                z_seg, z_lines, p->grid.nz);
  p->xcfg = xc;
  p->scfg[1] = yc;
  p->scfg[0] = zc;
  return HS2_OK;
}

int hs2_plan_set_option(hs2_plan *p, int option, int value) {
  if (!p) return fail(HS2_E_ARG, "hs2_plan_set_option: null plan");
  switch (option) {
    case HS2_OPT_COMBINED_X:
      p->use_comb = value != 0;
      return HS2_OK;
    default:
      return fail(HS2_E_ARG, "hs2_plan_set_option: unknown option %d", option);
  }
}

int hs2_plan_expand(hs2_plan *p, double *d_M, double *d_w, double *d_gz, double *d_gy,
                    double *d_gx, void *stream) {
  if (!p) return fail(HS2_E_ARG, "hs2_plan_expand: null plan");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_plan_expand: coefficients not set");
  if (p->mode == MODE_ASYM)
    return fail(HS2_E_UNSUPPORTED, "hs2_plan_expand: the plan's coefficients are the caller's "
                                   "(hs2_plan_set_dense_asym)");
  hipStream_t s = (hipStream_t)stream;
  if (p->mode == MODE_CLASS)
    HIP_TRY(hs2::launch_expand(grid_of(p), p->cv, d_M, d_w, d_gz, d_gy, d_gx, s));
  else
    HIP_TRY(hs2::launch_expand(grid_of(p), p->dv, d_M, d_w, d_gz, d_gy, d_gx, s));
  return HS2_OK;
}

int hs2_step(hs2_plan *p, const double *d_Tin, double *d_Tout, const uint8_t *d_vmap,
             const double *h_vtab, int nvtab, const double *d_sfield, double *d_surf,
             void *stream) {
  if (!p || !d_Tin || !d_Tout) return fail(HS2_E_ARG, "hs2_step: null argument");
  if (!aligned16(d_Tin) || !aligned16(d_Tout) || (d_sfield && !aligned16(d_sfield)))
    return fail(HS2_E_ARG, "hs2_step: field pointers must be 16-byte aligned");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_step: coefficients not set");
  if (d_surf && p->grid.nz < 2) return fail(HS2_E_ARG, "hs2_step: surface readout needs nz >= 2");
  if ((d_vmap != nullptr) != (h_vtab != nullptr))
    return fail(HS2_E_ARG, "hs2_step: class map and class table go together");
  if (d_vmap && (nvtab < 1 || nvtab > hs2::kMaxVtab))
    return fail(HS2_E_ARG, "hs2_step: nvtab=%d must be in 1..%d", nvtab, hs2::kMaxVtab);
  hipStream_t s = (hipStream_t)stream;
  if (d_vmap) {
    // a source step is rare: wait for earlier users of the table, then copy
    double tab[hs2::kMaxVtab];
    std::memset(tab, 0, sizeof tab);
    std::memcpy(tab, h_vtab, sizeof(double) * (size_t)nvtab);
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(p->d_vtab, tab, sizeof tab, hipMemcpyHostToDevice));
  }
  return run_step(p, d_Tin, d_Tout, d_vmap, d_vmap ? p->d_vtab : nullptr, d_sfield, d_surf, s);
}

int hs2_run(hs2_plan *p, double *d_T, int nsteps, const uint8_t *d_vmap,
            const double *h_vtab_steps, int nvtab, double *d_surf_hist, void *stream) {
  if (!p || !d_T) return fail(HS2_E_ARG, "hs2_run: null argument");
  if (!aligned16(d_T)) return fail(HS2_E_ARG, "hs2_run: field pointer must be 16-byte aligned");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_run: coefficients not set");
  if (nsteps < 0) return fail(HS2_E_ARG, "hs2_run: nsteps=%d", nsteps);
  if (d_surf_hist && p->grid.nz < 2) return fail(HS2_E_ARG, "hs2_run: surface readout needs nz >= 2");
  if (h_vtab_steps && (!d_vmap || nvtab < 1 || nvtab > hs2::kMaxVtab))
    return fail(HS2_E_ARG, "hs2_run: per-step source tables need the class map and 1..256 classes");
  hipStream_t s = (hipStream_t)stream;
  std::vector<int> active_row(nsteps, -1);
  std::vector<double> rows;
  if (h_vtab_steps) {
    for (int t = 0; t < nsteps; t++) {
      const double *row = h_vtab_steps + (size_t)t * nvtab;
      bool any = false;
      for (int v = 0; v < nvtab; v++) any |= row[v] != 0.0;
      if (!any) continue;
      active_row[t] = (int)(rows.size() / hs2::kMaxVtab);
      rows.resize(rows.size() + hs2::kMaxVtab, 0.0);
      std::memcpy(&rows[rows.size() - hs2::kMaxVtab], row, sizeof(double) * (size_t)nvtab);
    }
  }
  if (!rows.empty()) {
    HIP_TRY(hipStreamSynchronize(s));
    if (rows.size() > p->vsteps_cap) {
      (void)hipFree(p->d_vsteps);
      p->d_vsteps = nullptr;
      p->vsteps_cap = 0;
      HIP_TRY(hipMalloc(&p->d_vsteps, sizeof(double) * rows.size()));
      p->vsteps_cap = rows.size();
    }
    HIP_TRY(hipMemcpy(p->d_vsteps, rows.data(), sizeof(double) * rows.size(),
                      hipMemcpyHostToDevice));
  }
  const size_t sxy = (size_t)p->grid.ny * p->grid.nx;
  for (int t = 0; t < nsteps; t++) {
    const bool src = active_row[t] >= 0;
    const int rc = run_step(p, d_T, d_T, src ? d_vmap : nullptr,
                            src ? p->d_vsteps + (size_t)active_row[t] * hs2::kMaxVtab : nullptr,
                            nullptr, d_surf_hist ? d_surf_hist + (size_t)t * sxy : nullptr, s);
    if (rc) return rc;
  }
  return HS2_OK;
}

int hs2_run_timed(hs2_plan *p, double *d_T, int nsteps, int method, double *kernel_ms,
                  void *stream) {
  if (!p || !d_T || !kernel_ms || nsteps < 1 || !aligned16(d_T) || method < 0 || method > 2)
    return fail(HS2_E_ARG, "hs2_run_timed: bad argument");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_run_timed: coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  const bool ry = use_ry(p, false);
  // the three substep kernels of one source-free step; `extra` launches kernel
  // `twice` a second time first, into a scratch field (same inputs)
  auto step = [&](int twice, double *extra) -> hipError_t {
    hipError_t err = hipSuccess;
    if (twice == 0) err = sweep_x(p, d_T, extra, nullptr, nullptr, nullptr, s, ry);
    if (err == hipSuccess) err = sweep_x(p, d_T, p->d_ws, nullptr, nullptr, nullptr, s, ry);
    if (err == hipSuccess && twice == 1) err = sweep_s(p, 1, d_T, p->d_ws, extra, nullptr, s, ry);
    if (err == hipSuccess) err = sweep_s(p, 1, d_T, p->d_ws, p->d_ws, nullptr, s, ry);
    if (err == hipSuccess && twice == 2) err = sweep_s(p, 0, d_T, p->d_ws, extra, nullptr, s);
    if (err == hipSuccess) err = sweep_s(p, 0, d_T, p->d_ws, d_T, nullptr, s);
    return err;
  };
  const unsigned flags = method == 0 ? hipEventDefault : hipEventDisableSystemFence;
  const int nev = method == 2 ? 8 : nsteps * 4;
  std::vector<hipEvent_t> ev((size_t)nev, nullptr);
  double *extra = nullptr;
  int rc = HS2_OK;
  for (auto &e : ev)
    if (rc == HS2_OK && hipEventCreateWithFlags(&e, flags) != hipSuccess)
      rc = fail(HS2_E_HIP, "hs2_run_timed: hipEventCreateWithFlags");
  if (rc == HS2_OK && method == 2 && hipMalloc(&extra, sizeof(double) * (size_t)p->n) != hipSuccess)
    rc = fail(HS2_E_ALLOC, "hs2_run_timed: scratch field");
  hipError_t err = hipSuccess;
  if (rc == HS2_OK && method < 2) {
    // events around every substep kernel
    for (int t = 0; t < nsteps && err == hipSuccess; t++) {
      hipEvent_t *e = &ev[(size_t)t * 4];
      err = hipEventRecord(e[0], s);
      if (err == hipSuccess) err = sweep_x(p, d_T, p->d_ws, nullptr, nullptr, nullptr, s, ry);
      if (err == hipSuccess) err = hipEventRecord(e[1], s);
      if (err == hipSuccess) err = sweep_s(p, 1, d_T, p->d_ws, p->d_ws, nullptr, s, ry);
      if (err == hipSuccess) err = hipEventRecord(e[2], s);
      if (err == hipSuccess) err = sweep_s(p, 0, d_T, p->d_ws, d_T, nullptr, s);
      if (err == hipSuccess) err = hipEventRecord(e[3], s);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    double acc[4] = {0, 0, 0, 0};
    for (int t = 0; t < nsteps && err == hipSuccess; t++)
      for (int k = 0; k < 3 && err == hipSuccess; k++) {
        float ms = 0.f;
        err = hipEventElapsedTime(&ms, ev[(size_t)t * 4 + k], ev[(size_t)t * 4 + k + 1]);
        acc[k] += ms;
      }
    float all = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&all, ev.front(), ev.back());
    for (int k = 0; k < 3; k++) kernel_ms[k] = acc[k] / nsteps;
    kernel_ms[3] = (double)all / nsteps;
  } else if (rc == HS2_OK) {
    // one event pair around each of four runs: plain, and x, y or z twice
    float ms[4] = {0, 0, 0, 0};
    for (int r = 0; r < 4 && err == hipSuccess; r++) {
      err = hipEventRecord(ev[2 * r], s);
      for (int t = 0; t < nsteps && err == hipSuccess; t++) err = step(r - 1, extra);
      if (err == hipSuccess) err = hipEventRecord(ev[2 * r + 1], s);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    for (int r = 0; r < 4 && err == hipSuccess; r++)
      err = hipEventElapsedTime(&ms[r], ev[2 * r], ev[2 * r + 1]);
    for (int k = 0; k < 3; k++) kernel_ms[k] = (double)(ms[k + 1] - ms[0]) / nsteps;
    kernel_ms[3] = (double)ms[0] / nsteps;
  }
  if (rc == HS2_OK && err != hipSuccess)
    rc = fail(HS2_E_HIP, "hs2_run_timed: %s", hipGetErrorString(err));
  (void)hipFree(extra);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

size_t hs2_grad_work_doubles(const hs2_plan *p, int ckpt_every) {
  if (!p || ckpt_every < 1) return 0;
  return (3 * (size_t)ckpt_every + 6) * (size_t)p->n;
}

int hs2_grad_forward(hs2_plan *p, double *d_T, int nsteps, int ckpt_every, double *d_ckpt,
                     const double *d_src, const int32_t *h_src_step, int nsrc,
                     double *d_surf_hist, void *stream) {
  if (!p || !d_T || !d_ckpt || nsteps < 1 || ckpt_every < 1 || !aligned16(d_T) ||
      !aligned16(d_ckpt))
    return fail(HS2_E_ARG, "hs2_grad_forward: bad argument");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_grad_forward: coefficients not set");
  if (p->mode == MODE_ASYM)
    return fail(HS2_E_UNSUPPORTED, "hs2_grad_forward: no adjoint for per-side face coefficients");
  if (d_surf_hist && p->grid.nz < 2)
    return fail(HS2_E_ARG, "hs2_grad_forward: the surface readout needs nz >= 2");
  std::vector<int> src;
  int rc = source_schedule(nsteps, h_src_step, nsrc, d_src, &src, "hs2_grad_forward");
  if (rc != HS2_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)p->n, nyx = (size_t)p->grid.ny * p->grid.nx;
  for (int t = 0; t < nsteps; t++) {
    if (t % ckpt_every == 0)
      HIP_TRY(hipMemcpyAsync(d_ckpt + (size_t)(t / ckpt_every) * n, d_T, n * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    const double *sf = src[t] >= 0 ? d_src + (size_t)src[t] * n : nullptr;
    rc = run_step(p, d_T, d_T, nullptr, nullptr, sf, d_surf_hist ? d_surf_hist + t * nyx : nullptr, s);
    if (rc != HS2_OK) return rc;
  }
  return HS2_OK;
}

int hs2_grad_backward(hs2_plan *p, int nsteps, int ckpt_every, const double *d_ckpt,
                      const double *d_src, const int32_t *h_src_step, int nsrc,
                      const double *d_grad_hist, double *d_gM, double *d_gz, double *d_gy,
                      double *d_gx, double *d_gsrc, double *d_gT0, double *d_work,
                      size_t work_doubles, void *stream) {
  if (!p || !d_ckpt || !d_grad_hist || !d_gM || !d_gz || !d_gy || !d_gx || !d_work ||
      nsteps < 1 || ckpt_every < 1 || (nsrc > 0 && !d_gsrc))
    return fail(HS2_E_ARG, "hs2_grad_backward: bad argument");
  if (p->mode == MODE_NONE) return fail(HS2_E_STATE, "hs2_grad_backward: coefficients not set");
  if (p->mode == MODE_ASYM)
    return fail(HS2_E_UNSUPPORTED, "hs2_grad_backward: no adjoint for per-side face coefficients");
  if (p->grid.nz < 2) return fail(HS2_E_ARG, "hs2_grad_backward: the surface readout needs nz >= 2");
  if (!aligned16(d_ckpt) || !aligned16(d_work))
    return fail(HS2_E_ARG, "hs2_grad_backward: d_ckpt and d_work must be 16-byte aligned");
  const size_t need = hs2_grad_work_doubles(p, ckpt_every);
  if (work_doubles < need)
    return fail(HS2_E_ARG, "hs2_grad_backward: work buffer has %zu doubles, needs %zu",
                work_doubles, need);
  const hs2::Grid g = grid_of(p);
  if (!hs2::adj_line_config(g.nx).seg || !hs2::adj_line_config(g.ny).seg ||
      !hs2::adj_line_config(g.nz).seg)
    return fail(HS2_E_UNSUPPORTED, "hs2_grad_backward: axis length above 1024");
  std::vector<int> src;
  int rc = source_schedule(nsteps, h_src_step, nsrc, d_src, &src, "hs2_grad_backward");
  if (rc != HS2_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)p->n, nyx = (size_t)g.ny * g.nx;
  const int K = ckpt_every;
  // segment store: T^{n0} .. T^{n0+K} (K+1 fields), and T*, T** of each of
  // its K steps, so every state of the segment is computed once
  double *seg = d_work;
  double *segx = seg + (size_t)(K + 1) * n, *segy = segx + (size_t)K * n;
  double *ux = segy + (size_t)K * n, *uy = ux + n, *uz = uy + n;
  double *lam = uz + n, *lam2 = lam + n;
  HIP_TRY(hs2::launch_fill(d_gM, n, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gz, n, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gy, n, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gx, n, 0.0, s));
  HIP_TRY(hs2::launch_lam_from_readout(g, d_grad_hist + (size_t)(nsteps - 1) * nyx, lam, s));
  const int nseg = (nsteps + K - 1) / K;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int n0 = sg * K, n1 = std::min(nsteps, n0 + K);
    HIP_TRY(hipMemcpyAsync(seg, d_ckpt + (size_t)sg * n, n * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
    for (int t = n0; t < n1; t++) {  // recompute the segment's states
      const size_t q = (size_t)(t - n0);
      const double *sf = src[t] >= 0 ? d_src + (size_t)src[t] * n : nullptr;
      rc = step_states(p, seg + q * n, segx + q * n, segy + q * n, seg + (q + 1) * n, sf,
                       nullptr, s);
      if (rc != HS2_OK) return rc;
    }
    for (int t = n1 - 1; t >= n0; t--) {
      const size_t q = (size_t)(t - n0);
      HIP_TRY(adj_line(p, 0, lam, 0, uz, s));
      HIP_TRY(adj_line(p, 1, uz, 1, uy, s));
      HIP_TRY(adj_line(p, 2, uy, 1, ux, s));
      hs2::AdjFields f{};
      f.Tn = seg + q * n; f.Tx = segx + q * n; f.Ty = segy + q * n; f.Tn1 = seg + (q + 1) * n;
      f.ux = ux; f.uy = uy; f.uz = uz;
      f.ghist = t >= 1 ? d_grad_hist + (size_t)(t - 1) * nyx : nullptr;
      f.lam_out = lam2;
      f.gM = d_gM;
      f.gg[0] = d_gz; f.gg[1] = d_gy; f.gg[2] = d_gx;
      f.gsrc = src[t] >= 0 ? d_gsrc + (size_t)src[t] * n : nullptr;
      HIP_TRY(adj_stencil(p, f, s));
      std::swap(lam, lam2);
    }
  }
  if (d_gT0)
    HIP_TRY(hipMemcpyAsync(d_gT0, lam, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return HS2_OK;
}

int hs2_batch_create(const hs2_batch_inst *in, int ninst, int ckpt_every, hs2_batch **out) {
  if (!in || !out || ninst < 1 || ckpt_every < 2)
    return fail(HS2_E_ARG, "hs2_batch_create: bad argument");
  auto *b = new (std::nothrow) hs2_batch();
  if (!b) return fail(HS2_E_ALLOC, "hs2_batch_create: out of host memory");
  b->ninst = ninst;
  b->K = (ckpt_every + 1) & ~1;  // even: every segment starts on the same adjoint buffer
  long long off = 0, hoff = 0, ck = 0;
  for (int i = 0; i < ninst; i++) {
    const hs2_batch_inst &s = in[i];
    if (s.nz < 2 || s.ny < 1 || s.nx < 1 || s.nz > hs2::kBatchMaxLen ||
        s.ny > hs2::kBatchMaxLen || s.nx > hs2::kBatchMaxLen || s.nsteps < 1 ||
        s.src_step >= s.nsteps || !(s.dz > 0)) {
      delete b;
      return fail(HS2_E_ARG, "hs2_batch_create: instance %d: grid %dx%dx%d, nsteps %d, source "
                  "step %d (axes 1..%d, nz >= 2)", i, s.nz, s.ny, s.nx, s.nsteps, s.src_step,
                  hs2::kBatchMaxLen);
    }
    hs2::BatchInst I{};
    I.nz = s.nz; I.ny = s.ny; I.nx = s.nx; I.nsteps = s.nsteps;
    I.src_step = s.src_step; I.dz = s.dz;
    I.xseg = hs2::batch_x_seg(s.nx);
    I.cw[0] = hs2::batch_flat_cw(s.nz);
    I.cw[1] = hs2::batch_flat_cw(s.ny);
    I.cw[2] = hs2::batch_line_cw(s.nx, hs2::kBatchXThreads);
    const long long n = (long long)s.nz * s.ny * s.nx;
    I.off = off;
    I.hoff = hoff;
    I.ckoff = ck;
    off += (n + hs2::kBatchAlign - 1) / hs2::kBatchAlign * hs2::kBatchAlign;
    hoff += (long long)s.nsteps * s.ny * s.nx;
    ck += (long long)((s.nsteps + b->K - 1) / b->K) * n;
    b->maxsteps = std::max(b->maxsteps, s.nsteps);
    b->inst.push_back(I);
  }
  b->N = off;
  b->H = hoff;
  b->CK = ck;
  int rc = HS2_OK;
  auto bad = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == HS2_OK)
      rc = fail(HS2_E_HIP, "hs2_batch_create: %s: %s", what, hipGetErrorString(e));
  };
  bad(hipGetDevice(&b->device), "hipGetDevice");
  bad(hipMalloc(&b->d_inst, sizeof(hs2::BatchInst) * ninst), "hipMalloc");
  bad(hipMalloc(&b->d_ctr, 2 * sizeof(int)), "hipMalloc");
  if (rc == HS2_OK)
    bad(hipMemcpy(b->d_inst, b->inst.data(), sizeof(hs2::BatchInst) * ninst,
                  hipMemcpyHostToDevice), "hipMemcpy");
  for (int k = 0; k < NKIND && rc == HS2_OK; k++) {
    std::vector<int> b0(ninst + 1, 0), owner;
    for (int i = 0; i < ninst; i++) {
      const int nb = blocks_of(b->inst[i], k);
      b0[i + 1] = b0[i] + nb;
      owner.insert(owner.end(), nb, i);
    }
    b->nblk[k] = b0[ninst];
    bad(hipMalloc(&b->d_blk0[k], sizeof(int) * (ninst + 1)), "hipMalloc");
    bad(hipMalloc(&b->d_blk_inst[k], sizeof(int) * owner.size()), "hipMalloc");
    if (rc == HS2_OK) {
      bad(hipMemcpy(b->d_blk0[k], b0.data(), sizeof(int) * (ninst + 1), hipMemcpyHostToDevice),
          "hipMemcpy");
      bad(hipMemcpy(b->d_blk_inst[k], owner.data(), sizeof(int) * owner.size(),
                    hipMemcpyHostToDevice), "hipMemcpy");
    }
  }
  if (rc == HS2_OK) bad(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
  if (rc == HS2_OK) bad(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming), "event");
  if (rc == HS2_OK) bad(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming), "event");
  if (rc != HS2_OK) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return HS2_OK;
}

void hs2_batch_destroy(hs2_batch *b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  batch_free(b);
  delete b;
}

int hs2_batch_layout(const hs2_batch *b, int64_t *cell_off, int64_t *hist_off, int64_t *sizes) {
  if (!b) return fail(HS2_E_ARG, "hs2_batch_layout: null batch");
  for (int i = 0; i < b->ninst; i++) {
    if (cell_off) cell_off[i] = b->inst[i].off;
    if (hist_off) hist_off[i] = b->inst[i].hoff;
  }
  if (cell_off) cell_off[b->ninst] = b->N;
  if (hist_off) hist_off[b->ninst] = b->H;
  if (sizes) {
    sizes[0] = b->N;
    sizes[1] = b->H;
    sizes[2] = batch_work_doubles(b);
    sizes[3] = b->K;
    sizes[4] = b->maxsteps;
  }
  return HS2_OK;
}

int hs2_batch_set_dense(hs2_batch *b, const double *d_M, const uint8_t *d_w, const double *d_gz,
                        const double *d_gy, const double *d_gx) {
  if (!b || !d_M || !d_w || !d_gz || !d_gy || !d_gx)
    return fail(HS2_E_ARG, "hs2_batch_set_dense: null argument");
  if (!aligned16(d_M) || !aligned16(d_w) || !aligned16(d_gz) || !aligned16(d_gy) ||
      !aligned16(d_gx))
    return fail(HS2_E_ARG, "hs2_batch_set_dense: arrays must be 16-byte aligned");
  b->dv.M = d_M;
  b->dv.w = d_w;
  b->dv.g[0] = d_gz;
  b->dv.g[1] = d_gy;
  b->dv.g[2] = d_gx;
  b->have_coef = true;
  return HS2_OK;
}

int hs2_batch_set_options(hs2_batch *b, int use_graph, int chunk) {
  if (!b || chunk < 1) return fail(HS2_E_ARG, "hs2_batch_set_options: bad argument");
  b->use_graph = use_graph ? 1 : 0;
  b->chunk = chunk;
  return HS2_OK;
}

int hs2_batch_forward(hs2_batch *b, double *d_T, const double *d_src, double *d_hist,
                      double *d_work, size_t work_doubles, void *stream) {
  if (!b || !d_T || !d_hist || !d_work) return fail(HS2_E_ARG, "hs2_batch_forward: null argument");
  if (!b->have_coef) return fail(HS2_E_STATE, "hs2_batch_forward: coefficients not set");
  if (!aligned16(d_T) || !aligned16(d_work) || (d_src && !aligned16(d_src)))
    return fail(HS2_E_ARG, "hs2_batch_forward: fields must be 16-byte aligned");
  if (work_doubles < (size_t)batch_work_doubles(b))
    return fail(HS2_E_ARG, "hs2_batch_forward: work buffer too small");
  const BatchWork w = batch_work(b, d_work);
  HIP_TRY(hipEventRecord(b->ev_in, (hipStream_t)stream));
  HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  HIP_TRY(hs2::launch_batch_counter(b->d_ctr, 0, 0, 0, b->stream));
  const int C = std::min(b->chunk, b->maxsteps);
  const int reps = (b->maxsteps + C - 1) / C;
  const std::vector<const void *> key = {d_T, d_src, d_hist, d_work, b->dv.M, b->dv.w,
                                         b->dv.g[0], b->dv.g[1], b->dv.g[2],
                                         (const void *)(intptr_t)C};
  const int rc = batch_replay(b, &b->fwd_exec, &b->fwd_key, key, reps,
                              [&](hipStream_t s) { return emit_forward(b, d_T, d_src, d_hist, w, C, s); },
                              "hs2_batch_forward");
  if (rc != HS2_OK) return rc;
  HIP_TRY(hipEventRecord(b->ev_out, b->stream));
  HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
  return HS2_OK;
}

int hs2_batch_backward(hs2_batch *b, const double *d_src, const double *d_grad_hist,
                       double *d_gM, double *d_gz, double *d_gy, double *d_gx, double *d_gsrc,
                       double *d_work, size_t work_doubles, void *stream) {
  if (!b || !d_grad_hist || !d_gM || !d_gz || !d_gy || !d_gx || !d_work)
    return fail(HS2_E_ARG, "hs2_batch_backward: null argument");
  if (!b->have_coef) return fail(HS2_E_STATE, "hs2_batch_backward: coefficients not set");
  if (work_doubles < (size_t)batch_work_doubles(b))
    return fail(HS2_E_ARG, "hs2_batch_backward: work buffer too small");
  const BatchWork w = batch_work(b, d_work);
  hipStream_t s = b->stream;
  HIP_TRY(hipEventRecord(b->ev_in, (hipStream_t)stream));
  HIP_TRY(hipStreamWaitEvent(s, b->ev_in, 0));
  HIP_TRY(hs2::launch_fill(d_gM, b->N, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gz, b->N, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gy, b->N, 0.0, s));
  HIP_TRY(hs2::launch_fill(d_gx, b->N, 0.0, s));
  if (d_gsrc) HIP_TRY(hs2::launch_fill(d_gsrc, b->N, 0.0, s));
  HIP_TRY(hs2::launch_batch_lam_init(batch_map(b, KIND_CELL, hs2::RULE_ALL, 0), b->nblk[KIND_CELL],
                                     d_grad_hist, w.lam, s));
  HIP_TRY(hs2::launch_batch_counter(b->d_ctr, 1, 0, 0, s));
  const int reps = (b->maxsteps + b->K - 1) / b->K;
  const std::vector<const void *> key = {d_src, d_grad_hist, d_gM, d_gz, d_gy, d_gx, d_gsrc,
                                         d_work, b->dv.M, b->dv.w, b->dv.g[0], b->dv.g[1],
                                         b->dv.g[2]};
  const int rc = batch_replay(
      b, &b->bwd_exec, &b->bwd_key, key, reps,
      [&](hipStream_t st) {
        return emit_backward_segment(b, d_src, d_grad_hist, d_gM, d_gz, d_gy, d_gx, d_gsrc, w, st);
      },
      "hs2_batch_backward");
  if (rc != HS2_OK) return rc;
  HIP_TRY(hipEventRecord(b->ev_out, s));
  HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
  return HS2_OK;
}

int hs2_surface_temperature(const double *d_T, int nz, int ny, int nx, double dz, double *d_out,
                            void *stream) {
  if (!d_T || !d_out) return fail(HS2_E_ARG, "hs2_surface_temperature: null argument");
  if (nz < 2 || ny < 1 || nx < 1)
    return fail(HS2_E_ARG, "hs2_surface_temperature: needs nz >= 2 (got %dx%dx%d)", nz, ny, nx);
  HIP_TRY(hs2::launch_surface(d_T, ny, nx, dz, d_out, (hipStream_t)stream));
  return HS2_OK;
}

}  // extern "C"
