// hs2_step_noalloc: one ADI step that allocates nothing (declared in
// include/heatsim2_amd.h).  Its workspace (T*, then T** in place) is a
// file-scope device array of fixed size that every GPU the library runs on
// holds until the process ends; calls that overlap on one GPU share it and
// race.  Every workspace cell is written by the x sweep before it is read,
// so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "../../include/heatsim2_amd.h"
#include "hs2_launch.h"

namespace {

constexpr size_t kScratchCells = size_t(1) << 21;  // 16 MiB of doubles per device
constexpr int kScratchDevices = 64;

static __device__ __align__(256) double scratch[kScratchCells];

// Host code never uses `scratch` as an address: each device's copy sits at
// its own address, looked up once (the device's first call, with that
// device current) and passed to the kernels from then on.
std::mutex g_scratch_mu;
double *g_scratch_ptr[kScratchDevices] = {};

int scratch_of_current_device(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return hs2::fail(HS2_E_HIP, "hs2_step_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kScratchDevices)
    return hs2::fail(HS2_E_UNSUPPORTED, "hs2_step_noalloc: device %d beyond %d", dev,
                     kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return hs2::fail(HS2_E_HIP, "hs2_step_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                       dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return HS2_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

extern "C" int hs2_step_noalloc(const hs2_grid_desc *grid, const double *d_M, const uint8_t *d_w,
                                const double *d_gz, const double *d_gy, const double *d_gx,
                                const double *d_Tin, double *d_Tout, const double *d_sfield,
                                double *d_surf, void *stream) {
  // every check comes before the first HIP call
  if (!grid || !d_M || !d_w || !d_gz || !d_gy || !d_gx || !d_Tin || !d_Tout)
    return hs2::fail(HS2_E_ARG, "hs2_step_noalloc: null argument");
  const hs2_grid_desc &G = *grid;
  if (G.nz < 1 || G.ny < 1 || G.nx < 1)
    return hs2::fail(HS2_E_ARG, "hs2_step_noalloc: grid %dx%dx%d must be positive", G.nz, G.ny,
                     G.nx);
  const long long n = (long long)G.nz * G.ny * G.nx;
  if (n > (long long)kScratchCells)
    return hs2::fail(HS2_E_UNSUPPORTED,
                     "hs2_step_noalloc: grid of %lld cells exceeds the %zu-cell workspace of the "
                     "allocation-free step (use hs2_step)", n, kScratchCells);
  const hs2::XConfig xc = hs2::x_config_for(G.nx, G.nz);
  hs2::SConfig zc{0, 0}, yc{0, 0};
  if (!xc.cpl || !hs2::default_strided(G.nz, &zc) || !hs2::default_strided(G.ny, &yc))
    return hs2::fail(HS2_E_UNSUPPORTED, "hs2_step_noalloc: an axis of %dx%dx%d exceeds 1024",
                     G.nz, G.ny, G.nx);
  if (d_surf && (G.nz < 2 || !(G.dz > 0)))
    return hs2::fail(HS2_E_ARG, "hs2_step_noalloc: the surface readout needs nz >= 2, dz > 0");
  const void *ptrs[] = {d_M, d_w, d_gz, d_gy, d_gx, d_Tin, d_Tout, d_sfield, d_surf};
  for (const void *q : ptrs)
    if (q && !aligned16(q))
      return hs2::fail(HS2_E_ARG, "hs2_step_noalloc: arrays must be 16-byte aligned");
  double *ws = nullptr;
  const int rc = scratch_of_current_device(&ws);
  if (rc != HS2_OK) return rc;
  hs2::DenseView dv{};
  dv.M = d_M;
  dv.w = d_w;
  dv.g[0] = d_gz;
  dv.g[1] = d_gy;
  dv.g[2] = d_gx;
  const hs2::Grid g{G.nz, G.ny, G.nx};
  hipStream_t s = (hipStream_t)stream;
  // the kernels and launch shapes of hs2_step on a plan that was given these
  // coefficients with hs2_plan_set_dense
  hipError_t e = hs2::launch_sweep_x(xc, g, dv, d_Tin, ws, nullptr, nullptr, d_sfield, s);
  if (e == hipSuccess) e = hs2::launch_sweep_strided(yc, g, dv, 1, d_Tin, ws, ws, nullptr, G.dz, s);
  if (e == hipSuccess) e = hs2::launch_sweep_strided(zc, g, dv, 0, d_Tin, ws, d_Tout, d_surf, G.dz, s);
  if (e != hipSuccess)
    return hs2::fail(HS2_E_HIP, "hs2_step_noalloc: kernel launch: %s", hipGetErrorString(e));
  return HS2_OK;
}
