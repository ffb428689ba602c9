// Host-side launchers of the sweep kernels (defined in hs2_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hs2_device.h"

namespace hs2 {

// cells per lane for an x-line of nx cells (0 = unsupported, nx > 1024)
int x_cpl_for(int nx);

hipError_t launch_sweep_x(int cpl, const Grid &g, const Compact &cv, const double *Tn, double *Ts,
                          const uint8_t *vmap, const double *vtab, const double *sfield,
                          hipStream_t s);

size_t strided_lds_bytes(const Compact &cv, int P, int CW);

hipError_t launch_sweep_strided(int seg, int cw, const Grid &g, const Compact &cv, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s);

hipError_t launch_expand(const Grid &g, const Compact &cv, double *gz, double *gy, double *gx,
                         double *M, uint8_t *cflag, hipStream_t s);

hipError_t launch_surface(const double *T, int ny, int nx, double dz, double *out, hipStream_t s);

}  // namespace hs2
