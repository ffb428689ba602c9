// Adjoint of the Douglas ADI step (gfx950, FP64).
//
// Forward step, per cell (M = rho*c/dt, w = 1 COMPUTE / 0 FIXED, L_a the
// weighted face Laplacian of axis a, L_a = W G_a with G_a symmetric):
//   x: A_x T*      = (M + L_x/2 + L_y + L_z) T^n + w S      A_a = M - L_a/2
//   y: A_y T**     = M T*  - L_y/2 T^n
//   z: A_z T^{n+1} = M T** - L_z/2 T^n
// Given lam = dJ/dT^{n+1}, the adjoint runs the transposed solves in reverse
// axis order,
//   A_z^T u_z = lam,   A_y^T u_y = M u_z,   A_x^T u_x = M u_y,
// and then the adjoint stencil
//   dJ/dT^n = M u_x + G_x(w u_x / 2) + G_y(w (u_x - u_y / 2)) + G_z(w (u_x - u_z / 2))
// with the parameter sensitivities of this step (D(v) = v_c - v_m across the
// minus face of c on that axis, E(u) = w_m u_m - w_c u_c):
//   dJ/dM_c   -= u_x (T* - T^n) + u_y (T** - T*) + u_z (T^{n+1} - T**)
//   dJ/dg_x   += D(T* + T^n) E(u_x) / 2
//   dJ/dg_y   += D(T^n) E(u_x) + D(T** - T^n) E(u_y) / 2
//   dJ/dg_z   += D(T^n) E(u_x) + D(T^{n+1} - T^n) E(u_z) / 2
//   dJ/dS_c    = w u_x
// Row t of A_a^T along a line: diagonal M_t + w_t (g_t + g_{t+1}) / 2,
// sub-diagonal -w_{t-1} g_t / 2, super-diagonal -w_{t+1} g_{t+1} / 2 (g_t the
// conductance of the face between t-1 and t).  A_a is strictly row
// diagonally dominant, so A_a^T is strictly column dominant and the same
// segment elimination + PCR solves it without pivoting.
//
// No upstream counterpart: heatsim2 has no gradients.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hs2_adjoint.h"
#include "hs2_batch.h"

namespace hs2 {

// ------------------------------------------------- transposed line solve
// Solves A_a^T U = (scale_m ? M R : R) along every line of axis AXIS.
// Block (bx, by).  Axis 2 (x): line l = bx*CW + cl of the nz*ny x-lines.
// Axes 1, 0: columns i = bx*CW + cl, by = k (y) or j (z).  A block has P
// thread segments of SEG cells per line; P = blockDim.x / CW.  U must not
// alias R.
// FLAT (batched, axes 0 and 1): columns as in sweep_strided_body's FLAT, by
// = P segments per line.
template <class V, int SEG, int AXIS, bool FLAT = false>
__device__ __forceinline__ void adj_line_body(Grid g, V view, int CW, const double *R,
                                              int scale_m, double *U, int bx, int by) {
  static_assert(!(FLAT && AXIS == 2), "x lines are not flattened");
  extern __shared__ double lds[];
  view.load_tables(lds, AXIS);
  const int P = FLAT ? by : blockDim.x / CW;
  double *ra = lds + ((view.lds_doubles(AXIS) + 1) & ~1);
  double *rc = ra + 2 * P * CW;
  double *rr = rc + 2 * P * CW;
  __syncthreads();

  const int cl = threadIdx.x % CW, p = threadIdx.x / CW;
  const bool act = !FLAT || p < P;
  int L, stride, base;
  bool valid;
  if constexpr (FLAT) {
    L = AXIS == 1 ? g.ny : g.nz;
    stride = AXIS == 1 ? g.nx : g.nx * g.ny;
    const int ncol = g.nx * (AXIS == 1 ? g.nz : g.ny);
    const int col = bx * CW + cl;
    valid = act && col < ncol;
    const int cc = min(col, ncol - 1);
    if (AXIS == 1) {
      const int k = cc / g.nx;
      base = k * g.ny * g.nx + (cc - k * g.nx);
    } else {
      base = cc;
    }
  } else if (AXIS == 2) {
    const int nl = g.nz * g.ny, line = bx * CW + cl;
    valid = line < nl;
    L = g.nx;
    stride = 1;
    base = (valid ? line : nl - 1) * g.nx;
  } else {
    const int i = bx * CW + cl;
    valid = i < g.nx;
    const int ic = valid ? i : g.nx - 1;
    L = AXIS == 1 ? g.ny : g.nz;
    stride = AXIS == 1 ? g.nx : g.nx * g.ny;
    base = AXIS == 1 ? by * g.ny * g.nx + ic : by * g.nx + ic;
  }
  const int t0 = p * SEG;

  // entries of cells t0-1 .. t0+SEG (clamped into the line) and the RHS
  typename V::Raw raw[SEG + 2];
  double Rv[SEG];
#pragma unroll
  for (int q = 0; q < SEG + 2; q++)
    raw[q] = view.fetch(AXIS, base + min(max(t0 + q - 1, 0), L - 1) * stride);
#pragma unroll
  for (int q = 0; q < SEG; q++) Rv[q] = R[base + min(t0 + q, L - 1) * stride];

  double a[SEG], c[SEG], r[SEG];
  Entry em = view.resolve(lds, raw[0]);
  Entry ec = view.resolve(lds, raw[1]);
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    const Entry ep = view.resolve(lds, raw[q + 2]);
    const double gm = t >= 1 ? ec.g : 0.0;
    const double gp = t + 1 <= L - 1 ? ep.g : 0.0;
    const double b = ec.M + 0.5 * ec.w * (gm + gp);
    const double inv = 1.0 / b;
    const double rhs = scale_m ? ec.M * Rv[q] : Rv[q];
    const bool ok = valid && t < L;
    a[q] = ok ? -0.5 * em.w * gm * inv : 0.0;
    c[q] = ok ? -0.5 * ep.w * gp * inv : 0.0;
    r[q] = ok ? rhs * inv : 0.0;
    em = ec;
    ec = ep;
  }
  double As, Cs, Rs, Ae, Ce, Re;
  segment_reduce<SEG>(a, c, r, As, Cs, Rs, Ae, Ce, Re);
  if (act) {
    ra[(2 * p) * CW + cl] = As; rc[(2 * p) * CW + cl] = Cs; rr[(2 * p) * CW + cl] = Rs;
    ra[(2 * p + 1) * CW + cl] = Ae; rc[(2 * p + 1) * CW + cl] = Ce; rr[(2 * p + 1) * CW + cl] = Re;
  }
  __syncthreads();
  lds_pcr_ends(ra, rc, rr, p, P, CW, cl, act);
  if (!valid) return;
  const double xs = rr[(2 * p) * CW + cl], xe = rr[(2 * p + 1) * CW + cl];
  double x[SEG];
  x[0] = xs;
  x[SEG - 1] = xe;
#pragma unroll
  for (int q = 1; q < SEG - 1; q++) x[q] = r[q] - a[q] * xs - c[q] * xe;
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    if (t < L) U[base + t * stride] = x[q];
  }
}

// adj_line_body for x lines of per-voxel coefficients, staged through LDS:
// the block's CW x-lines are consecutive in memory, so the right-hand side is
// loaded cell by cell along the lines (coalesced) into padded LDS rows, and
// the solution leaves through LDS the same way; each thread reads its
// segment's x entries (M, g, w) from memory itself (staging them as well
// quadrupled the block's LDS and cost more in occupancy than the strided
// loads do).  LDS: 6 P CW doubles of reduced rows, then CW rows of (L | 1)
// doubles.
template <int SEG>
__device__ __forceinline__ void adj_x_staged_body(Grid g, DenseView view, int CW, const double *R,
                                                  int scale_m, double *U, int bx) {
  extern __shared__ double lds[];
  const int P = blockDim.x / CW;
  const int L = g.nx, RW = g.nx | 1;
  double *ra = lds;
  double *rc = ra + 2 * P * CW;
  double *rr = rc + 2 * P * CW;
  double *sR = rr + 2 * P * CW;
  const int nl = g.nz * g.ny, l0 = bx * CW;
  const int nlb = min(CW, nl - l0);
  const int ncell = nlb * L;
  const long long c0 = (long long)l0 * L;
  // UN cells per thread per batch, all loads issued before the first use
  constexpr int UN = 8;
  for (int e0 = 0; e0 < ncell; e0 += UN * (int)blockDim.x) {
    double vr[UN];
#pragma unroll
    for (int u = 0; u < UN; u++) {
      const long long c = c0 + min(e0 + u * (int)blockDim.x + (int)threadIdx.x, ncell - 1);
      vr[u] = R[c];
    }
#pragma unroll
    for (int u = 0; u < UN; u++) {
      const int e = e0 + u * (int)blockDim.x + (int)threadIdx.x;
      if (e < ncell) {
        const int lc = e / L;
        sR[lc * RW + (e - lc * L)] = vr[u];
      }
    }
  }
  __syncthreads();
  const int cl = threadIdx.x % CW, p = threadIdx.x / CW;
  const bool valid = cl < nlb;
  const int row = (valid ? cl : 0) * RW;
  const long long grow = c0 + (long long)(valid ? cl : 0) * L;
  const int t0 = p * SEG;
  auto entry = [&](int t) {
    const long long o = grow + min(max(t, 0), L - 1);
    return Entry{view.M[o], (double)view.w[o], view.g[2][o]};
  };
  double a[SEG], c[SEG], r[SEG];
  Entry em = entry(t0 - 1);
  Entry ec = entry(t0);
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    const Entry ep = entry(t + 1);
    const double gm = t >= 1 ? ec.g : 0.0;
    const double gp = t + 1 <= L - 1 ? ep.g : 0.0;
    const double b = ec.M + 0.5 * ec.w * (gm + gp);
    const double inv = 1.0 / b;
    const double Rq = sR[row + min(t, L - 1)];
    const double rhs = scale_m ? ec.M * Rq : Rq;
    const bool ok = valid && t < L;
    a[q] = ok ? -0.5 * em.w * gm * inv : 0.0;
    c[q] = ok ? -0.5 * ep.w * gp * inv : 0.0;
    r[q] = ok ? rhs * inv : 0.0;
    em = ec;
    ec = ep;
  }
  double As, Cs, Rs, Ae, Ce, Re;
  segment_reduce<SEG>(a, c, r, As, Cs, Rs, Ae, Ce, Re);
  ra[(2 * p) * CW + cl] = As; rc[(2 * p) * CW + cl] = Cs; rr[(2 * p) * CW + cl] = Rs;
  ra[(2 * p + 1) * CW + cl] = Ae; rc[(2 * p + 1) * CW + cl] = Ce; rr[(2 * p + 1) * CW + cl] = Re;
  __syncthreads();
  // (every thread has built its rows from sR before the barriers in here,
  // so sR can take the solution afterwards)
  lds_pcr_ends(ra, rc, rr, p, P, CW, cl);
  const double xs = rr[(2 * p) * CW + cl], xe = rr[(2 * p + 1) * CW + cl];
  if (valid) {
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int t = t0 + q;
      const double x = q == 0 ? xs : q == SEG - 1 ? xe : r[q] - a[q] * xs - c[q] * xe;
      if (t < L) sR[row + t] = x;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ncell; e += blockDim.x) {
    const int lc = e / L;
    U[c0 + e] = sR[lc * RW + (e - lc * L)];
  }
}

template <class V, int SEG, int AXIS>
__global__ __launch_bounds__(256) void k_adj_line(Grid g, V view, int CW, const double *R,
                                                  int scale_m, double *U) {
  const int nb = gridDim.x * gridDim.y;
  const int r = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, nb);
  adj_line_body<V, SEG, AXIS>(g, view, CW, R, scale_m, U, r % gridDim.x, r / gridDim.x);
}

// ------------------------------------------------------ adjoint stencil
// dJ/dT^n of cell c = (k, j, i) (plus the readout adjoint of this step's input
// field when ghist != null: the insulating-top readout T0 - 5/16 (T1 - T0) of
// T^n), returned; accumulates the cell's sensitivities into gM and gg[axis]
// (the cell's minus face on that axis), and writes gsrc = w u_x when gsrc !=
// null.  `lds`: the view's tables (all axes).
template <class V>
__device__ __forceinline__ double adj_stencil_cell(Grid g, const V &view, const double *lds,
                                                   const AdjFields &f, int c, int k, int j,
                                                   int i) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int sxy = nx * ny;
  const int coord[3] = {k, j, i};
  const int len[3] = {nz, ny, nx};
  const int step[3] = {sxy, nx, 1};

  const Entry e0 = view.resolve(view.table(lds, 0), view.fetch(0, c));
  const double Mc = e0.M, wc = e0.w;
  const double uxc = f.ux[c], uyc = f.uy[c], uzc = f.uz[c];
  const double Tnc = f.Tn[c], Txc = f.Tx[c], Tyc = f.Ty[c], T1c = f.Tn1[c];

  double lam = Mc * uxc;
  // v^a: x: w u_x / 2, y: w (u_x - u_y / 2), z: w (u_x - u_z / 2)
  auto vof = [](int a, double w, double ux, double uy, double uz) {
    return a == 2 ? 0.5 * w * ux : a == 1 ? w * (ux - 0.5 * uy) : w * (ux - 0.5 * uz);
  };
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double vc = vof(a, wc, uxc, uyc, uzc);
    if (coord[a] > 0) {
      const int m = c - step[a];
      const Entry ec = view.resolve(view.table(lds, a), view.fetch(a, c));
      const Entry emn = view.resolve(view.table(lds, a), view.fetch(a, m));
      const double gf = ec.g, wm = emn.w;
      const double uxm = f.ux[m], uym = f.uy[m], uzm = f.uz[m];
      lam += gf * (vof(a, wm, uxm, uym, uzm) - vc);
      const double Tnm = f.Tn[m];
      const double Ex = wm * uxm - wc * uxc;
      double gg;
      if (a == 2) {
        gg = 0.5 * ((Txc + Tnc) - (f.Tx[m] + Tnm)) * Ex;
      } else if (a == 1) {
        const double Ey = wm * uym - wc * uyc;
        gg = (Tnc - Tnm) * Ex + 0.5 * ((Tyc - Tnc) - (f.Ty[m] - Tnm)) * Ey;
      } else {
        const double Ez = wm * uzm - wc * uzc;
        gg = (Tnc - Tnm) * Ex + 0.5 * ((T1c - Tnc) - (f.Tn1[m] - Tnm)) * Ez;
      }
      f.gg[a][c] += gg;
    }
    if (coord[a] < len[a] - 1) {
      const int pn = c + step[a];
      const Entry ep = view.resolve(view.table(lds, a), view.fetch(a, pn));
      lam += ep.g * (vof(a, ep.w, f.ux[pn], f.uy[pn], f.uz[pn]) - vc);
    }
  }
  if (f.ghist) {
    const int rem = c - k * sxy;
    if (k == 0) lam += (21.0 / 16.0) * f.ghist[rem];
    else if (k == 1) lam -= (5.0 / 16.0) * f.ghist[rem];
  }
  f.gM[c] -= uxc * (Txc - Tnc) + uyc * (Tyc - Txc) + uzc * (T1c - Tyc);
  if (f.gsrc) f.gsrc[c] = wc * uxc;
  return lam;
}

// One thread per cell: lam_out = adj_stencil_cell.
template <class V>
__device__ __forceinline__ void adj_stencil_body(Grid g, V view, const AdjFields &f, int blk) {
  extern __shared__ double lds[];
  view.load_tables(lds, -1);
  __syncthreads();
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int sxy = nx * ny, n = sxy * nz;
  const int c = blk * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int k = c / sxy, rem = c - k * sxy, j = rem / nx, i = rem - j * nx;
  f.lam_out[c] = adj_stencil_cell<V>(g, view, lds, f, c, k, j, i);
}

template <class V>
__global__ __launch_bounds__(256) void k_adj_stencil(Grid g, V view, AdjFields f) {
  adj_stencil_body<V>(g, view, f, blockIdx.x);
}

// lam = readout adjoint of ghist (ny, nx) on the top two layers, 0 elsewhere
__global__ void k_lam_from_readout(Grid g, const double *ghist, double *lam) {
  const int sxy = g.nx * g.ny, n = sxy * g.nz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int k = c / sxy, rem = c - k * sxy;
    lam[c] = k == 0 ? (21.0 / 16.0) * ghist[rem] : k == 1 ? -(5.0 / 16.0) * ghist[rem] : 0.0;
  }
}

__global__ void k_fill(double *a, size_t n, double v) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    a[t] = v;
}

// ---------------------------------------------------------- launchers
AdjLineConfig adj_line_config(int L) {
  AdjLineConfig c;
  c.seg = L <= 64 ? 4 : (L <= 512 ? 8 : 16);
  const int P = (L + c.seg - 1) / c.seg;
  c.cw = 64;
  while (c.cw > 1 && P * c.cw > 256) c.cw >>= 1;
  if (P * c.cw > 256) c.seg = 0;  // unsupported length
  return c;
}

template <class V, int SEG, int AXIS>
static hipError_t launch_adj_line_t(const Grid &g, const V &v, int CW, const double *R,
                                    int scale_m, double *U, hipStream_t s) {
  const int L = AXIS == 2 ? g.nx : AXIS == 1 ? g.ny : g.nz;
  const int P = (L + SEG - 1) / SEG;
  dim3 grid;
  if (AXIS == 2) grid = dim3((g.nz * g.ny + CW - 1) / CW, 1);
  else grid = dim3((g.nx + CW - 1) / CW, AXIS == 1 ? g.nz : g.ny);
  const size_t lds = sizeof(double) * ((size_t)((v.lds_doubles(AXIS) + 1) & ~1) + 6 * (size_t)P * CW);
  hipLaunchKernelGGL((k_adj_line<V, SEG, AXIS>), grid, dim3(P * CW), lds, s, g, v, CW, R, scale_m,
                     U);
  return hipGetLastError();
}

template <class V, int SEG>
static hipError_t launch_adj_line_s(const Grid &g, const V &v, int axis, int CW, const double *R,
                                    int scale_m, double *U, hipStream_t s) {
  switch (axis) {
    case 0: return launch_adj_line_t<V, SEG, 0>(g, v, CW, R, scale_m, U, s);
    case 1: return launch_adj_line_t<V, SEG, 1>(g, v, CW, R, scale_m, U, s);
    case 2: return launch_adj_line_t<V, SEG, 2>(g, v, CW, R, scale_m, U, s);
    default: return hipErrorInvalidValue;
  }
}

template <class V>
static hipError_t launch_adj_line_v(const Grid &g, const V &v, int axis, const double *R,
                                    int scale_m, double *U, hipStream_t s) {
  const int L = axis == 2 ? g.nx : axis == 1 ? g.ny : g.nz;
  const AdjLineConfig c = adj_line_config(L);
  switch (c.seg) {
    case 4: return launch_adj_line_s<V, 4>(g, v, axis, c.cw, R, scale_m, U, s);
    case 8: return launch_adj_line_s<V, 8>(g, v, axis, c.cw, R, scale_m, U, s);
    case 16: return launch_adj_line_s<V, 16>(g, v, axis, c.cw, R, scale_m, U, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_adj_line(const Grid &g, const ClassView &v, int axis, const double *R,
                           int scale_m, double *U, hipStream_t s) {
  return launch_adj_line_v(g, v, axis, R, scale_m, U, s);
}
hipError_t launch_adj_line(const Grid &g, const DenseView &v, int axis, const double *R,
                           int scale_m, double *U, hipStream_t s) {
  return launch_adj_line_v(g, v, axis, R, scale_m, U, s);
}

template <class V>
static hipError_t launch_adj_stencil_t(const Grid &g, const V &v, const AdjFields &f,
                                       hipStream_t s) {
  const long n = (long)g.nx * g.ny * g.nz;
  const size_t lds = sizeof(double) * (size_t)v.lds_doubles(-1);
  hipLaunchKernelGGL((k_adj_stencil<V>), dim3((unsigned)((n + 255) / 256)), dim3(256), lds, s, g, v,
                     f);
  return hipGetLastError();
}
hipError_t launch_adj_stencil(const Grid &g, const ClassView &v, const AdjFields &f,
                              hipStream_t s) {
  return launch_adj_stencil_t(g, v, f, s);
}
hipError_t launch_adj_stencil(const Grid &g, const DenseView &v, const AdjFields &f,
                              hipStream_t s) {
  return launch_adj_stencil_t(g, v, f, s);
}

hipError_t launch_lam_from_readout(const Grid &g, const double *ghist, double *lam,
                                   hipStream_t s) {
  hipLaunchKernelGGL(k_lam_from_readout, dim3(1024), dim3(256), 0, s, g, ghist, lam);
  return hipGetLastError();
}

hipError_t launch_fill(double *a, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, a, n, v);
  return hipGetLastError();
}


// ------------------------------------------------------ batched adjoint
template <int SEG, int AXIS>
__global__ __launch_bounds__(kBatchThreads) void kb_adj_line(BatchMap m, DenseView v,
                                                             const double *R, int scale_m,
                                                             double *U) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;
  const int cw = I.cw[AXIS];
  if constexpr (AXIS == 2)
    adj_x_staged_body<SEG>(Grid{I.nz, I.ny, I.nx}, offset_view(v, I.off), cw, R + I.off, scale_m,
                           U + I.off, local);
  else
    adj_line_body<DenseView, SEG, AXIS, true>(Grid{I.nz, I.ny, I.nx}, offset_view(v, I.off), cw,
                                              R + I.off, scale_m, U + I.off, local,
                                              batch_line_segs(AXIS == 1 ? I.ny : I.nz));
}

// f: base pointers of the concatenated fields; f.ghist: base of the
// concatenated dJ/dhist; f.gsrc: base of the concatenated source gradients.
__global__ __launch_bounds__(256) void kb_adj_stencil(BatchMap m, DenseView v, AdjFields f) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;
  const long long o = I.off;
  AdjFields g = f;
  g.Tn += o; g.Tx += o; g.Ty += o; g.Tn1 += o;
  g.ux += o; g.uy += o; g.uz += o;
  g.lam_out += o; g.gM += o;
  g.gg[0] += o; g.gg[1] += o; g.gg[2] += o;
  g.ghist = (f.ghist && t >= 1) ? f.ghist + I.hoff + (long long)(t - 1) * I.ny * I.nx : nullptr;
  g.gsrc = (f.gsrc && t == I.src_step) ? f.gsrc + o : nullptr;
  adj_stencil_body<DenseView>(Grid{I.nz, I.ny, I.nx}, offset_view(v, o), g, local);
}

// backward segment start: checkpoint of the segment -> segment slot K - (e - s)
__global__ void kb_ckpt_load(BatchMap m, const double *ck, double *seg, long long N) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;  // RULE_SEG: t = segment start s
  const long long n = (long long)I.nz * I.ny * I.nx;
  const long long c = (long long)local * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int e = I.nsteps - m.ctr[1] * m.K;
  const int slot = m.K - (e - t);
  ck += I.ckoff + (long long)m.ctr[1] * n;
  seg[(long long)slot * N + I.off + c] = ck[c];
}

// lam = readout adjoint of the last history row of each instance
__global__ void kb_lam_init(BatchMap m, const double *ghist, double *lam) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;
  const long long nyx = (long long)I.ny * I.nx, n = nyx * I.nz;
  const long long c = (long long)local * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const long long k = c / nyx, rem = c - k * nyx;
  const double *gh = ghist + I.hoff + (long long)(I.nsteps - 1) * nyx;
  lam[I.off + c] = k == 0 ? (21.0 / 16.0) * gh[rem] : k == 1 ? -(5.0 / 16.0) * gh[rem] : 0.0;
}

// ctr[which] = value (set) or += value (add); one thread
__global__ void kb_counter(int *ctr, int which, int value, int add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[which] = add ? ctr[which] + value : value;
}

hipError_t launch_batch_adj_line(const BatchMap &m, int nblocks, int axis, const DenseView &v,
                                 const double *R, int scale_m, double *U, hipStream_t s) {
  // x lines: blocks of kBatchXThreads staging CW rows of (L | 1) right-hand
  // sides, CW L <= kBatchXThreads kBatchSeg
  const int threads = axis == 2 ? kBatchXThreads : kBatchThreads;
  const size_t rows = (size_t)kBatchXThreads * kBatchSeg + kBatchXThreads;
  const size_t shm = sizeof(double) * (6 * (size_t)threads + (axis == 2 ? rows : 0));
  const dim3 grid(nblocks), block(threads);
  switch (axis) {
    case 0: hipLaunchKernelGGL((kb_adj_line<kBatchSeg, 0>), grid, block, shm, s, m, v, R, scale_m, U); break;
    case 1: hipLaunchKernelGGL((kb_adj_line<kBatchSeg, 1>), grid, block, shm, s, m, v, R, scale_m, U); break;
    default: hipLaunchKernelGGL((kb_adj_line<kBatchSeg, 2>), grid, block, shm, s, m, v, R, scale_m, U); break;
  }
  return hipGetLastError();
}

hipError_t launch_batch_adj_stencil(const BatchMap &m, int nblocks, const DenseView &v,
                                    const AdjFields &f, hipStream_t s) {
  hipLaunchKernelGGL(kb_adj_stencil, dim3(nblocks), dim3(256), 0, s, m, v, f);
  return hipGetLastError();
}

hipError_t launch_batch_ckpt_load(const BatchMap &m, int nblocks, const double *ck, double *seg,
                                  long long N, hipStream_t s) {
  hipLaunchKernelGGL(kb_ckpt_load, dim3(nblocks), dim3(256), 0, s, m, ck, seg, N);
  return hipGetLastError();
}

hipError_t launch_batch_lam_init(const BatchMap &m, int nblocks, const double *ghist, double *lam,
                                 hipStream_t s) {
  hipLaunchKernelGGL(kb_lam_init, dim3(nblocks), dim3(256), 0, s, m, ghist, lam);
  return hipGetLastError();
}

hipError_t launch_batch_counter(int *ctr, int which, int value, int add, hipStream_t s) {
  hipLaunchKernelGGL(kb_counter, dim3(1), dim3(64), 0, s, ctr, which, value, add);
  return hipGetLastError();
}

}  // namespace hs2
