// Host-side launchers of the sweep kernels (defined in hs2_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hs2_device.h"

namespace hs2 {

struct XConfig {  // x sweep: CPL cells per lane, G lanes per line, LPB lines per
  int cpl, g, lpb;  // block of the two-phase kernel (0: single-phase kernel);
  int ty = 0;       // two-phase: the block's lines as LPB/TY planes x TY rows (0: in order)
};
struct SConfig {  // strided sweep: SEG cells per thread, CW lines per block
  int seg, cw;
};

XConfig x_config_for(int nx, int nz);  // {0,0}: unsupported (nx > 1024)
// untuned y / z shape for lines of L cells; 0 when L > 1024
int default_strided(int L, SConfig *cfg);
// record a message for hs2_last_error() on this thread; returns code
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// ry: the x sweep writes y's right-hand side M T* - (Ly/2) T^n instead of T*
// and the y sweep reads it (class maps, steps without sources, x shapes for
// which x_ry_ok holds); the z sweep is unchanged
bool x_ry_ok(XConfig c, const Grid &g);

hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const ClassView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry = false);
// the combined class plane (full-line source-free steps, x_ry_ok(cfg, g))
hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const CombView &v, const double *Tn,
                          double *Ts, hipStream_t s, bool ry);
hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const DenseView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry = false);
hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const AsymView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry = false);

size_t strided_lds_bytes(const ClassView &v, int axis, int P, int CW);
size_t strided_lds_bytes(const DenseView &v, int axis, int P, int CW);

hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const ClassView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry = false);
hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const DenseView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry = false);
hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const AsymView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry = false);

// K0 of the combined plane: mark the (z, y, x) class triples that occur in a
// 2^24-bit map, then give each cell the rank of its triple among them
hipError_t launch_comb_mark(int n, const uint8_t *cz, const uint8_t *cy, const uint8_t *cx,
                            uint32_t *bits, hipStream_t s);
hipError_t launch_comb_assign(int n, const uint8_t *cz, const uint8_t *cy, const uint8_t *cx,
                              const uint32_t *bits, const int *rank0, uint8_t *comb,
                              hipStream_t s);
hipError_t launch_classify(const Grid &g, const uint8_t *mat, const uint8_t *bz, const uint8_t *by,
                           const uint8_t *bx, const uint8_t *lut, const uint8_t *edge, int nmat,
                           int lutlen, uint8_t *cz, uint8_t *cy, uint8_t *cx, hipStream_t s);

hipError_t launch_expand(const Grid &g, const ClassView &v, double *M, double *w, double *gz,
                         double *gy, double *gx, hipStream_t s);
hipError_t launch_expand(const Grid &g, const DenseView &v, double *M, double *w, double *gz,
                         double *gy, double *gx, hipStream_t s);

hipError_t launch_surface(const double *T, int ny, int nx, double dz, double *out, hipStream_t s);

}  // namespace hs2
