// C ABI of heatsim2_amd (declared in include/heatsim2_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/heatsim2_amd.h"
#include "hs2_launch.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(HS2_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// LDS budget per block for the tables (doubles); beyond this the compact
// description is refused.
constexpr int kMaxTabDoubles = 4096;

}  // namespace

struct hs2_plan {
  hs2_grid_desc grid{};
  int n = 0;
  int device = 0;
  hs2::Compact cv{};
  double *d_tab = nullptr;
  bool ready = false;
  double *d_ws = nullptr;     // n doubles: T* then T**
  double *d_vtab = nullptr;   // kMaxVtab doubles (hs2_step)
  double *d_vsteps = nullptr; // per-step source tables (hs2_run)
  size_t vsteps_cap = 0;      // doubles
  int *d_check = nullptr;     // per-block maxima of the maps (validation)
  int cpl = 0;
  int seg[2] = {0, 0};        // strided config per axis index 0=z, 1=y
  int cw[2] = {0, 0};
};

namespace {

int strided_config(int L, int *seg, int *cw) {
  int s;
  if (L <= 4) s = 2;
  else if (L <= 16) s = 4;
  else if (L <= 64) s = 8;
  else if (L <= 256) s = 16;
  else if (L <= 1024) s = 32;
  else return 0;
  const int P = (L + s - 1) / s;
  int w = 64;
  while (w > 8 && P * w > 512) w >>= 1;
  if (P * w > 512) return 0;  // __launch_bounds__(512) of the strided kernel
  *seg = s;
  *cw = w;
  return 1;
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

constexpr int kCheckBlocks = 256;

// largest entry of a device uint8 map
int map_max(const uint8_t *d_map, size_t n, int *d_scratch, int *result) {
  hipLaunchKernelGGL(k_map_max, dim3(kCheckBlocks), dim3(256), 0, 0, d_map, n, d_scratch);
  HIP_TRY(hipGetLastError());
  int part[kCheckBlocks];
  HIP_TRY(hipMemcpy(part, d_scratch, sizeof part, hipMemcpyDeviceToHost));
  *result = *std::max_element(part, part + kCheckBlocks);
  return HS2_OK;
}

int run_step(hs2_plan *p, const double *Tin, double *Tout, const uint8_t *vmap,
             const double *d_vtab, const double *sfield, double *surf, hipStream_t s) {
  const hs2::Grid g{p->grid.nz, p->grid.ny, p->grid.nx};
  HIP_TRY(hs2::launch_sweep_x(p->cpl, g, p->cv, Tin, p->d_ws, vmap, d_vtab, sfield, s));
  HIP_TRY(hs2::launch_sweep_strided(p->seg[1], p->cw[1], g, p->cv, 1, Tin, p->d_ws, p->d_ws,
                                    nullptr, p->grid.dz, s));
  HIP_TRY(hs2::launch_sweep_strided(p->seg[0], p->cw[0], g, p->cv, 0, Tin, p->d_ws, Tout, surf,
                                    p->grid.dz, s));
  return HS2_OK;
}

}  // namespace

extern "C" {

const char *hs2_last_error(void) { return g_last_error.c_str(); }

const char *hs2_version(void) { return "heatsim2_amd 0.1 (gfx950, fp64)"; }

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
  const int cpl = hs2::x_cpl_for(G.nx);
  if (!cpl) return fail(HS2_E_UNSUPPORTED, "hs2_plan_create: nx=%d > 1024 unsupported", G.nx);
  hs2_plan *p = new hs2_plan();
  p->grid = G;
  p->n = (int)n;
  p->cpl = cpl;
  if (!strided_config(G.ny, &p->seg[1], &p->cw[1]) || !strided_config(G.nz, &p->seg[0], &p->cw[0])) {
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
  (void)hipFree(p->d_ws);
  (void)hipFree(p->d_vtab);
  (void)hipFree(p->d_vsteps);
  (void)hipFree(p->d_tab);
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
  const long long glen = 3LL * nb * nmat * nmat;
  const long long tab_len = glen + 2LL * nmat;
  if (tab_len > kMaxTabDoubles)
    return fail(HS2_E_UNSUPPORTED,
                "hs2_plan_set_compact: table of %lld doubles (nb=%d, nmat=%d) exceeds %d", tab_len,
                nb, nmat, kMaxTabDoubles);
  for (int m = 0; m < nmat; m++) {
    if (h_compute[m] > 1) return fail(HS2_E_ARG, "hs2_plan_set_compact: compute flag not 0/1");
    if (!(h_mtab[m] > 0.0))
      return fail(HS2_E_ARG, "hs2_plan_set_compact: M[%d]=%g must be positive", m, h_mtab[m]);
  }
  for (long long t = 0; t < glen; t++)
    if (!(h_gtab[t] >= 0.0) || h_gtab[t] > 1e300)
      return fail(HS2_E_ARG, "hs2_plan_set_compact: conductance table entry %lld = %g invalid", t,
                  h_gtab[t]);
  const hs2_grid_desc &G = p->grid;
  // every map entry must index the tables
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
  std::vector<double> tab((size_t)tab_len);
  std::memcpy(tab.data(), h_gtab, sizeof(double) * (size_t)glen);
  for (int m = 0; m < nmat; m++) {
    tab[(size_t)glen + m] = h_mtab[m];
    tab[(size_t)glen + nmat + m] = h_compute[m] ? 1.0 : 0.0;
  }
  // tables must fit the LDS of the strided kernels next to their reduced systems
  const hs2::Grid g{G.nz, G.ny, G.nx};
  hs2::Compact cv{d_mat, d_bz, d_by, d_bx, nullptr, nmat, nb, (int)tab_len};
  for (int ax = 0; ax < 2; ax++) {
    const int L = ax == 1 ? G.ny : G.nz;
    const int P = (L + p->seg[ax] - 1) / p->seg[ax];
    if (hs2::strided_lds_bytes(cv, P, p->cw[ax]) > 64 * 1024)
      return fail(HS2_E_UNSUPPORTED, "hs2_plan_set_compact: LDS budget exceeded");
  }
  (void)g;
  if (p->d_tab) (void)hipFree(p->d_tab);
  p->d_tab = nullptr;
  HIP_TRY(hipMalloc(&p->d_tab, sizeof(double) * tab.size()));
  HIP_TRY(hipMemcpy(p->d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  cv.tab = p->d_tab;
  p->cv = cv;
  p->ready = true;
  return HS2_OK;
}

int hs2_plan_expand(hs2_plan *p, double *d_gz, double *d_gy, double *d_gx, double *d_M,
                    uint8_t *d_cflag, void *stream) {
  if (!p) return fail(HS2_E_ARG, "hs2_plan_expand: null plan");
  if (!p->ready) return fail(HS2_E_STATE, "hs2_plan_expand: coefficients not set");
  const hs2::Grid g{p->grid.nz, p->grid.ny, p->grid.nx};
  HIP_TRY(hs2::launch_expand(g, p->cv, d_gz, d_gy, d_gx, d_M, d_cflag, (hipStream_t)stream));
  return HS2_OK;
}

int hs2_step(hs2_plan *p, const double *d_Tin, double *d_Tout, const uint8_t *d_vmap,
             const double *h_vtab, int nvtab, const double *d_sfield, double *d_surf,
             void *stream) {
  if (!p || !d_Tin || !d_Tout) return fail(HS2_E_ARG, "hs2_step: null argument");
  if (d_surf && p->grid.nz < 2) return fail(HS2_E_ARG, "hs2_step: surface readout needs nz >= 2");
  if (!p->ready) return fail(HS2_E_STATE, "hs2_step: coefficients not set");
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
  if (!p->ready) return fail(HS2_E_STATE, "hs2_run: coefficients not set");
  if (nsteps < 0) return fail(HS2_E_ARG, "hs2_run: nsteps=%d", nsteps);
  if (d_surf_hist && p->grid.nz < 2) return fail(HS2_E_ARG, "hs2_run: surface readout needs nz >= 2");
  if (h_vtab_steps && (!d_vmap || nvtab < 1 || nvtab > hs2::kMaxVtab))
    return fail(HS2_E_ARG, "hs2_run: per-step source tables need the class map and 1..256 classes");
  hipStream_t s = (hipStream_t)stream;
  // gather the steps that carry a source
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
    if (rows.size() > p->vsteps_cap) {
      HIP_TRY(hipStreamSynchronize(s));
      (void)hipFree(p->d_vsteps);
      p->d_vsteps = nullptr;
      p->vsteps_cap = 0;
      HIP_TRY(hipMalloc(&p->d_vsteps, sizeof(double) * rows.size()));
      p->vsteps_cap = rows.size();
    }
    HIP_TRY(hipMemcpyAsync(p->d_vsteps, rows.data(), sizeof(double) * rows.size(),
                           hipMemcpyHostToDevice, s));
    // the host vector must outlive the copy
    HIP_TRY(hipStreamSynchronize(s));
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

int hs2_run_timed(hs2_plan *p, double *d_T, int nsteps, double *kernel_ms, void *stream) {
  if (!p || !d_T || !kernel_ms || nsteps < 1) return fail(HS2_E_ARG, "hs2_run_timed: bad argument");
  if (!p->ready) return fail(HS2_E_STATE, "hs2_run_timed: coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  const hs2::Grid g{p->grid.nz, p->grid.ny, p->grid.nx};
  std::vector<hipEvent_t> ev((size_t)nsteps * 4, nullptr);
  int rc = HS2_OK;
  for (auto &e : ev) {
    if (hipEventCreate(&e) != hipSuccess) { rc = fail(HS2_E_HIP, "hs2_run_timed: hipEventCreate"); break; }
  }
  for (int t = 0; t < nsteps && rc == HS2_OK; t++) {
    hipEvent_t *e = &ev[(size_t)t * 4];
    hipError_t err = hipEventRecord(e[0], s);
    if (err == hipSuccess) err = hs2::launch_sweep_x(p->cpl, g, p->cv, d_T, p->d_ws, nullptr, nullptr, nullptr, s);
    if (err == hipSuccess) err = hipEventRecord(e[1], s);
    if (err == hipSuccess)
      err = hs2::launch_sweep_strided(p->seg[1], p->cw[1], g, p->cv, 1, d_T, p->d_ws, p->d_ws, nullptr, p->grid.dz, s);
    if (err == hipSuccess) err = hipEventRecord(e[2], s);
    if (err == hipSuccess)
      err = hs2::launch_sweep_strided(p->seg[0], p->cw[0], g, p->cv, 0, d_T, p->d_ws, d_T, nullptr, p->grid.dz, s);
    if (err == hipSuccess) err = hipEventRecord(e[3], s);
    if (err != hipSuccess) rc = fail(HS2_E_HIP, "hs2_run_timed: %s", hipGetErrorString(err));
  }
  if (rc == HS2_OK) {
    hipError_t err = hipEventSynchronize(ev.back());
    double acc[3] = {0, 0, 0};
    for (int t = 0; t < nsteps && err == hipSuccess; t++)
      for (int k = 0; k < 3 && err == hipSuccess; k++) {
        float ms = 0.f;
        err = hipEventElapsedTime(&ms, ev[(size_t)t * 4 + k], ev[(size_t)t * 4 + k + 1]);
        acc[k] += ms;
      }
    if (err != hipSuccess) rc = fail(HS2_E_HIP, "hs2_run_timed: %s", hipGetErrorString(err));
    for (int k = 0; k < 3; k++) kernel_ms[k] = acc[k] / nsteps;
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
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
