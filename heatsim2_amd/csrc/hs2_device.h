// Device-side building blocks shared by the heatsim2_amd kernels (gfx950).
//
// Coefficient model.  For a COMPUTE cell c the explicit operator on axis a is
//   (L_a w)_c = g_a-(c) (w_{c-e_a} - w_c) + g_a+(c) (w_{c+e_a} - w_c)
// and M_c = rho*c/dt; a FIXED cell has M = 1 and no couplings.  g is a
// property of the face (the same number from both COMPUTE sides), looked up
// in a per-plan table indexed by (face class, axis, minus material, plus
// material) -- the numbers the upstream setup derives symbolically per cell
// (crank_nicolson.pyx:298-386).  Outer faces never carry flux.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hs2 {

constexpr int kWave = 64;
constexpr int kMaxVtab = 256;

struct Grid {
  int nz, ny, nx;
};

// Compact description: device uint8 maps + a device copy of the tables.
// Table layout (doubles): g[nb][3][nmat][nmat], then M[nmat], then
// compute-flag[nmat] (1.0 / 0.0).
struct Compact {
  const uint8_t *__restrict__ mat;
  const uint8_t *__restrict__ bz;
  const uint8_t *__restrict__ by;
  const uint8_t *__restrict__ bx;
  const double *__restrict__ tab;
  int nmat, nb, tab_len;
};

// Logical block index that keeps consecutive logical blocks on one XCD.
// Hardware hands block b to XCD b % 8; give each XCD a contiguous range.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int x = b & 7, idx = b >> 3;
  return x * q + (x < r ? x : r) + idx;
}

__device__ __forceinline__ void load_table(double *lds, const Compact &cv) {
  for (int t = threadIdx.x; t < cv.tab_len; t += blockDim.x) lds[t] = cv.tab[t];
}

__device__ __forceinline__ double gface(const double *lds, const Compact &cv, int bcls, int axis,
                                        int mm, int mp) {
  return lds[((bcls * 3 + axis) * cv.nmat + mm) * cv.nmat + mp];
}

__device__ __forceinline__ double shfl_up_d(double v, int d) {
  return __shfl_up(v, (unsigned)d, kWave);
}
__device__ __forceinline__ double shfl_down_d(double v, int d) {
  return __shfl_down(v, (unsigned)d, kWave);
}

// One PCR reduction of a normalised row  a x_{i-1} + x_i + c x_{i+1} = r
// with its neighbour rows L (i-1) and R (i+1), giving a row that couples
// i-2 and i+2.  Absent neighbours are passed as zeros.
__device__ __forceinline__ void pcr_combine(double &a, double &c, double &r, double aL,
                                            double cL, double rL, double aR, double cR,
                                            double rR) {
  const double k = 1.0 / (1.0 - a * cL - c * aR);
  const double an = -a * aL * k;
  const double cn = -c * cR * k;
  const double rn = (r - a * rL - c * rR) * k;
  a = an;
  c = cn;
  r = rn;
}

// Segment elimination.  Rows q = 0..N-1 of one contiguous segment, already
// normalised (diagonal 1): a[q] x_{q-1} + x_q + c[q] x_{q+1} = r[q].
// Afterwards:
//   (As, Cs, Rs): As x_{before} + x_0     + Cs x_{N-1}  = Rs
//   (Ae, Ce, Re): Ae x_0        + x_{N-1} + Ce x_{after} = Re
// and, for interior rows 1..N-2, a[q] x_0 + x_q + c[q] x_{N-1} = r[q]
// (a, c, r overwritten).  N >= 2.
template <int N>
__device__ __forceinline__ void segment_reduce(double (&a)[N], double (&c)[N], double (&r)[N],
                                               double &As, double &Cs, double &Rs, double &Ae,
                                               double &Ce, double &Re) {
  static_assert(N >= 2, "segment needs two rows");
  // forward: rows 2..N-1 in terms of x_0 (row 1 already is)
#pragma unroll
  for (int q = 2; q < N; q++) {
    const double k = 1.0 / (1.0 - a[q] * c[q - 1]);
    const double an = -a[q] * a[q - 1] * k;
    c[q] = c[q] * k;
    r[q] = (r[q] - a[q] * r[q - 1]) * k;
    a[q] = an;
  }
  Ae = a[N - 1];
  Ce = c[N - 1];
  Re = r[N - 1];
  if constexpr (N == 2) {
    As = a[0];
    Cs = c[0];
    Rs = r[0];
  } else {
    // backward: rows N-3..1 in terms of x_0 and x_{N-1} (row N-2 already is)
#pragma unroll
    for (int q = N - 3; q >= 1; q--) {
      a[q] = a[q] - c[q] * a[q + 1];
      r[q] = r[q] - c[q] * r[q + 1];
      c[q] = -c[q] * c[q + 1];
    }
    const double k = 1.0 / (1.0 - c[0] * a[1]);
    As = a[0] * k;
    Cs = -c[0] * c[1] * k;
    Rs = (r[0] - c[0] * r[1]) * k;
  }
}

// readout of surface_temperature.py:31-36 for the first two z-layers.
__device__ __forceinline__ double surface_readout(double T0, double T1, double dz) {
  const double dz2 = dz * dz;
  const double a = (T1 - T0) / (2.0 * dz2);
  const double cc = T0 - .25 * a * dz2;
  const double dT = T1 - T0;
  const double lin = T0 - dT / 2.0;
  return (lin + cc) / 2.0;
}

}  // namespace hs2
