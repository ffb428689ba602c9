// Adjoint kernels of the ADI step (defined in hs2_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hs2_device.h"

namespace hs2 {

// Fields of one adjoint step.  Inputs: the forward states of the step
// (T^n, T*, T**, T^{n+1}), the transposed-solve results u_x, u_y, u_z, and
// optionally the readout adjoint ghist (ny, nx) of T^n.  Outputs: lam_out =
// dJ/dT^n, gM and gg[axis] (accumulated), gsrc = w u_x (written if non-null).
struct AdjFields {
  const double *Tn, *Tx, *Ty, *Tn1;
  const double *ux, *uy, *uz;
  const double *ghist;
  double *lam_out, *gM;
  double *gg[3];
  double *gsrc;
};

struct AdjLineConfig {
  int seg, cw;  // seg 0: unsupported line length
};

AdjLineConfig adj_line_config(int L);

// U = A_axis^{-T} (scale_m ? M R : R) along every line of `axis`
hipError_t launch_adj_line(const Grid &g, const ClassView &v, int axis, const double *R,
                           int scale_m, double *U, hipStream_t s);
hipError_t launch_adj_line(const Grid &g, const DenseView &v, int axis, const double *R,
                           int scale_m, double *U, hipStream_t s);
hipError_t launch_adj_stencil(const Grid &g, const ClassView &v, const AdjFields &f,
                              hipStream_t s);
hipError_t launch_adj_stencil(const Grid &g, const DenseView &v, const AdjFields &f,
                              hipStream_t s);
hipError_t launch_lam_from_readout(const Grid &g, const double *ghist, double *lam,
                                   hipStream_t s);
hipError_t launch_fill(double *a, size_t n, double v, hipStream_t s);

}  // namespace hs2
