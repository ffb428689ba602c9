// heatsim2_amd sweep kernels for gfx950 (MI355X), FP64.
//
// One Douglas ADI Crank-Nicolson step in delta form (algebraically the split
// of alternatingdirection_c_pyx.pyx:475-493):
//   x substep:  (M - Lx/2) T*    = (M + Lx/2 + Ly + Lz) T^n + D S
//   y substep:  (M - Ly/2) T**   = M T*  - (Ly/2) T^n
//   z substep:  (M - Lz/2) T^n+1 = M T** - (Lz/2) T^n
// Each substep is one kernel that forms the right-hand side and solves the
// line tridiagonal systems in the same pass (the upstream does a CSR product
// then tridiag.tridiagsolve, alternatingdirection_c_pyx.pyx:402-412).
//
// x (contiguous): one wavefront per x-line; lane l owns CPL consecutive
//   cells, eliminates them locally, and the 2x64-row reduced system is
//   solved by parallel cyclic reduction across the wavefront with lane
//   shuffles.
// y, z (strided): CW consecutive lines per block, coalesced along x; each
//   thread runs the Thomas-style elimination of one SEG-cell segment of one
//   line in registers, the 2P-row reduced system of each line is solved by
//   PCR in LDS, then each thread back-substitutes its segment.  With one
//   segment per line this is the plain one-thread-per-line Thomas sweep.
#include "hs2_device.h"
#include "hs2_launch.h"

namespace hs2 {

// ---------------------------------------------------------------- x sweep
template <int CPL>
__global__ __launch_bounds__(256) void k_sweep_x(Grid g, Compact cv,
                                                 const double *__restrict__ Tn,
                                                 double *__restrict__ Ts,
                                                 const uint8_t *__restrict__ vmap,
                                                 const double *__restrict__ vtab_g,
                                                 const double *__restrict__ sfield) {
  extern __shared__ double lds[];
  load_table(lds, cv);
  double *vtab = lds + cv.tab_len;
  if (vmap)
    for (int t = threadIdx.x; t < kMaxVtab; t += blockDim.x) vtab[t] = vtab_g[t];
  __syncthreads();

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nlines = nz * ny;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int line = b * 4 + wave;
  if (line >= nlines) return;  // whole wavefront leaves together
  const int k = line / ny, j = line - k * ny;
  const int sxy = nx * ny;
  const int base = line * nx;
  const int i0 = lane * CPL;
  const int gofs = cv.nb * 3 * cv.nmat * cv.nmat;

  double T[CPL];
  int mc[CPL];
#pragma unroll
  for (int q = 0; q < CPL; q++) {
    const int i = i0 + q;
    T[q] = i < nx ? Tn[base + i] : 0.0;
    mc[q] = i < nx ? cv.mat[base + i] : 0;
  }
  const double TL = shfl_up_d(T[CPL - 1], 1);
  const double TR = shfl_down_d(T[0], 1);
  const int mL = __shfl_up(mc[CPL - 1], 1u, kWave);
  const int mR = __shfl_down(mc[0], 1u, kWave);

  // conductances of faces i0 .. i0+CPL (face i lies between cells i-1, i)
  double gx[CPL + 1];
  const int frow = line * (nx + 1);
#pragma unroll
  for (int q = 0; q <= CPL; q++) {
    const int i = i0 + q;
    if (i >= 1 && i <= nx - 1) {
      const int mm = q == 0 ? mL : mc[q - 1];
      const int mp = q == CPL ? mR : mc[q];
      gx[q] = gface(lds, cv, cv.bx[frow + i], 2, mm, mp);
    } else {
      gx[q] = 0.0;
    }
  }

  double a[CPL], c[CPL], r[CPL];
#pragma unroll
  for (int q = 0; q < CPL; q++) {
    const int i = i0 + q;
    if (i >= nx) {
      a[q] = 0.0; c[q] = 0.0; r[q] = 0.0;
      continue;
    }
    const int cell = base + i;
    const int m = mc[q];
    const double Tc = T[q];
    if (lds[gofs + cv.nmat + m] == 0.0) {  // FIXED: T* = T^n
      a[q] = 0.0; c[q] = 0.0; r[q] = Tc;
      continue;
    }
    const double M = lds[gofs + m];
    double rhs = M * Tc;
    const double Txm = q > 0 ? T[q - 1] : TL;
    const double Txp = q < CPL - 1 ? T[q + 1] : TR;
    rhs += 0.5 * (gx[q] * (Txm - Tc) + gx[q + 1] * (Txp - Tc));
    if (j > 0) {
      const double gym = gface(lds, cv, cv.by[(k * (ny + 1) + j) * nx + i], 1, cv.mat[cell - nx], m);
      rhs += gym * (Tn[cell - nx] - Tc);
    }
    if (j < ny - 1) {
      const double gyp = gface(lds, cv, cv.by[(k * (ny + 1) + j + 1) * nx + i], 1, m, cv.mat[cell + nx]);
      rhs += gyp * (Tn[cell + nx] - Tc);
    }
    if (k > 0) {
      const double gzm = gface(lds, cv, cv.bz[cell], 0, cv.mat[cell - sxy], m);
      rhs += gzm * (Tn[cell - sxy] - Tc);
    }
    if (k < nz - 1) {
      const double gzp = gface(lds, cv, cv.bz[cell + sxy], 0, m, cv.mat[cell + sxy]);
      rhs += gzp * (Tn[cell + sxy] - Tc);
    }
    if (vmap) rhs += vtab[vmap[cell]];
    if (sfield) rhs += sfield[cell];
    const double am = -0.5 * gx[q], cp = -0.5 * gx[q + 1];
    const double inv = 1.0 / (M - am - cp);
    a[q] = am * inv;
    c[q] = cp * inv;
    r[q] = rhs * inv;
  }

  double x[CPL];
  if constexpr (CPL == 1) {
    double aa = a[0], cc = c[0], rr = r[0];
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      double aL = shfl_up_d(aa, d), cL = shfl_up_d(cc, d), rL = shfl_up_d(rr, d);
      double aR = shfl_down_d(aa, d), cR = shfl_down_d(cc, d), rR = shfl_down_d(rr, d);
      if (lane < d) { aL = 0.0; cL = 0.0; rL = 0.0; }
      if (lane + d >= kWave) { aR = 0.0; cR = 0.0; rR = 0.0; }
      pcr_combine(aa, cc, rr, aL, cL, rL, aR, cR, rR);
    }
    x[0] = rr;
  } else {
    double As, Cs, Rs, Ae, Ce, Re;
    segment_reduce<CPL>(a, c, r, As, Cs, Rs, Ae, Ce, Re);
    // reduced order s0 e0 s1 e1 ...: first PCR step pairs s_p with e_{p-1}, e_p
    // and e_p with s_p, s_{p+1}, which splits it into an s- and an e-system.
    {
      double aL = shfl_up_d(Ae, 1), cL = shfl_up_d(Ce, 1), rL = shfl_up_d(Re, 1);
      double aR = shfl_down_d(As, 1), cR = shfl_down_d(Cs, 1), rR = shfl_down_d(Rs, 1);
      if (lane == 0) { aL = 0.0; cL = 0.0; rL = 0.0; }
      if (lane == kWave - 1) { aR = 0.0; cR = 0.0; rR = 0.0; }
      double as = As, cs = Cs, rs = Rs, ae = Ae, ce = Ce, re = Re;
      pcr_combine(as, cs, rs, aL, cL, rL, Ae, Ce, Re);
      pcr_combine(ae, ce, re, As, Cs, Rs, aR, cR, rR);
      As = as; Cs = cs; Rs = rs; Ae = ae; Ce = ce; Re = re;
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      double aL = shfl_up_d(As, d), cL = shfl_up_d(Cs, d), rL = shfl_up_d(Rs, d);
      double aR = shfl_down_d(As, d), cR = shfl_down_d(Cs, d), rR = shfl_down_d(Rs, d);
      double eaL = shfl_up_d(Ae, d), ecL = shfl_up_d(Ce, d), erL = shfl_up_d(Re, d);
      double eaR = shfl_down_d(Ae, d), ecR = shfl_down_d(Ce, d), erR = shfl_down_d(Re, d);
      if (lane < d) { aL = cL = rL = 0.0; eaL = ecL = erL = 0.0; }
      if (lane + d >= kWave) { aR = cR = rR = 0.0; eaR = ecR = erR = 0.0; }
      pcr_combine(As, Cs, Rs, aL, cL, rL, aR, cR, rR);
      pcr_combine(Ae, Ce, Re, eaL, ecL, erL, eaR, ecR, erR);
    }
    x[0] = Rs;
    x[CPL - 1] = Re;
#pragma unroll
    for (int q = 1; q < CPL - 1; q++) x[q] = r[q] - a[q] * Rs - c[q] * Re;
  }
#pragma unroll
  for (int q = 0; q < CPL; q++) {
    const int i = i0 + q;
    if (i < nx) Ts[base + i] = x[q];
  }
}

// ---------------------------------------------------------- y / z sweeps
// axis 1 (y): columns (k, i), line stride nx, faces in by.
// axis 0 (z): columns (j, i), line stride nx*ny, faces in bz.
// Solves (M - La/2) X = M Y - (La/2) T^n along each line; X may alias Y or
// T^n (every load of a block happens before its first barrier and every
// store after its last, and a block owns whole lines).
template <int SEG>
__global__ __launch_bounds__(512) void k_sweep_strided(Grid g, Compact cv, int axis, int CW,
                                                        const double *Tn, const double *Y,
                                                        double *X, double *surf, double dz) {
  extern __shared__ double lds[];
  load_table(lds, cv);
  const int P = blockDim.x / CW;
  const int tabpad = (cv.tab_len + 1) & ~1;
  double *ra = lds + tabpad;
  double *rc = ra + 2 * P * CW;
  double *rr = rc + 2 * P * CW;
  __syncthreads();

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int cl = threadIdx.x % CW, p = threadIdx.x / CW;
  const int i = blockIdx.x * CW + cl;
  const int other = blockIdx.y;
  const bool valid = i < nx;
  int L, stride, base, fbase, fstride;
  const uint8_t *fmap;
  if (axis == 1) {
    L = ny; stride = nx; base = other * ny * nx + i;
    fmap = cv.by; fbase = other * (ny + 1) * nx + i; fstride = nx;
  } else {
    L = nz; stride = nx * ny; base = other * nx + i;
    fmap = cv.bz; fbase = base; fstride = nx * ny;
  }
  const int gofs = cv.nb * 3 * cv.nmat * cv.nmat;
  const int t0 = p * SEG;

  // T^n and materials over the segment plus one halo cell each side
  double T[SEG + 2];
  int m[SEG + 2];
#pragma unroll
  for (int q = 0; q < SEG + 2; q++) {
    const int t = t0 + q - 1;
    const bool in = valid && t >= 0 && t < L;
    T[q] = in ? Tn[base + t * stride] : 0.0;
    m[q] = in ? cv.mat[base + t * stride] : 0;
  }
  double gf[SEG + 1];
#pragma unroll
  for (int q = 0; q <= SEG; q++) {
    const int t = t0 + q;
    gf[q] = (valid && t >= 1 && t <= L - 1)
                ? gface(lds, cv, fmap[fbase + t * fstride], axis, m[q], m[q + 1])
                : 0.0;
  }
  double a[SEG], c[SEG], r[SEG];
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    if (!valid || t >= L) {
      a[q] = 0.0; c[q] = 0.0; r[q] = 0.0;
      continue;
    }
    const int mm = m[q + 1];
    const double y = Y[base + t * stride];
    if (lds[gofs + cv.nmat + mm] == 0.0) {
      a[q] = 0.0; c[q] = 0.0; r[q] = y;
      continue;
    }
    const double M = lds[gofs + mm];
    const double Tc = T[q + 1];
    const double rhs = M * y - 0.5 * (gf[q] * (T[q] - Tc) + gf[q + 1] * (T[q + 2] - Tc));
    const double am = -0.5 * gf[q], cp = -0.5 * gf[q + 1];
    const double inv = 1.0 / (M - am - cp);
    a[q] = am * inv;
    c[q] = cp * inv;
    r[q] = rhs * inv;
  }
  double As, Cs, Rs, Ae, Ce, Re;
  segment_reduce<SEG>(a, c, r, As, Cs, Rs, Ae, Ce, Re);
  const int R = 2 * P;
  ra[(2 * p) * CW + cl] = As; rc[(2 * p) * CW + cl] = Cs; rr[(2 * p) * CW + cl] = Rs;
  ra[(2 * p + 1) * CW + cl] = Ae; rc[(2 * p + 1) * CW + cl] = Ce; rr[(2 * p + 1) * CW + cl] = Re;
  __syncthreads();
  for (int d = 1; d < R; d <<= 1) {
    double na[2], nc[2], nr[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = 2 * p + h;
      double aa = ra[row * CW + cl], cc = rc[row * CW + cl], r0 = rr[row * CW + cl];
      double aL = 0.0, cL = 0.0, rL = 0.0, aR = 0.0, cR = 0.0, rR = 0.0;
      if (row - d >= 0) { aL = ra[(row - d) * CW + cl]; cL = rc[(row - d) * CW + cl]; rL = rr[(row - d) * CW + cl]; }
      if (row + d < R) { aR = ra[(row + d) * CW + cl]; cR = rc[(row + d) * CW + cl]; rR = rr[(row + d) * CW + cl]; }
      pcr_combine(aa, cc, r0, aL, cL, rL, aR, cR, rR);
      na[h] = aa; nc[h] = cc; nr[h] = r0;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int row = 2 * p + h;
      ra[row * CW + cl] = na[h]; rc[row * CW + cl] = nc[h]; rr[row * CW + cl] = nr[h];
    }
    __syncthreads();
  }
  const double xs = rr[(2 * p) * CW + cl], xe = rr[(2 * p + 1) * CW + cl];
  if (!valid) return;
  double x[SEG];
  x[0] = xs;
  x[SEG - 1] = xe;
#pragma unroll
  for (int q = 1; q < SEG - 1; q++) x[q] = r[q] - a[q] * xs - c[q] * xe;
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    if (t < L) X[base + t * stride] = x[q];
  }
  if (surf && axis == 0 && p == 0 && L >= 2) surf[base] = surface_readout(x[0], x[1], dz);
}

// --------------------------------------------------------------- helpers
__global__ void k_expand(Grid g, Compact cv, double *gz, double *gy, double *gx, double *M,
                         uint8_t *cflag) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int n = nx * ny * nz;
  const int gofs = cv.nb * 3 * cv.nmat * cv.nmat;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int k = c / (nx * ny), rem = c - k * nx * ny, j = rem / nx, i = rem - j * nx;
    const int m = cv.mat[c];
    if (M) M[c] = cv.tab[gofs + m];
    if (cflag) cflag[c] = cv.tab[gofs + cv.nmat + m] != 0.0;
    // minus faces of this cell, plus the far outer face when on the last layer
    if (gz) {
      gz[c] = k > 0 ? cv.tab[((cv.bz[c] * 3 + 0) * cv.nmat + cv.mat[c - nx * ny]) * cv.nmat + m] : 0.0;
      if (k == nz - 1) gz[c + nx * ny] = 0.0;
    }
    if (gy) {
      const int f = (k * (ny + 1) + j) * nx + i;
      gy[f] = j > 0 ? cv.tab[((cv.by[f] * 3 + 1) * cv.nmat + cv.mat[c - nx]) * cv.nmat + m] : 0.0;
      if (j == ny - 1) gy[f + nx] = 0.0;
    }
    if (gx) {
      const int f = (k * ny + j) * (nx + 1) + i;
      gx[f] = i > 0 ? cv.tab[((cv.bx[f] * 3 + 2) * cv.nmat + cv.mat[c - 1]) * cv.nmat + m] : 0.0;
      if (i == nx - 1) gx[f + 1] = 0.0;
    }
  }
}

__global__ void k_surface(const double *T, int ny, int nx, double dz, double *out) {
  const int sxy = nx * ny;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < sxy; c += gridDim.x * blockDim.x)
    out[c] = surface_readout(T[c], T[c + sxy], dz);
}

// ------------------------------------------------------------- launchers
template <int CPL>
static hipError_t launch_x_t(const Grid &g, const Compact &cv, const double *Tn, double *Ts,
                             const uint8_t *vmap, const double *vtab, const double *sfield,
                             hipStream_t s) {
  const int nlines = g.nz * g.ny;
  const int blocks = (nlines + 3) / 4;
  const size_t shm = sizeof(double) * (cv.tab_len + (vmap ? kMaxVtab : 0));
  hipLaunchKernelGGL(k_sweep_x<CPL>, dim3(blocks), dim3(256), shm, s, g, cv, Tn, Ts, vmap, vtab,
                     sfield);
  return hipGetLastError();
}

int x_cpl_for(int nx) {
  const int need = (nx + kWave - 1) / kWave;
  if (need <= 1) return 1;
  if (need <= 2) return 2;
  if (need <= 4) return 4;
  if (need <= 8) return 8;
  if (need <= 16) return 16;
  return 0;
}

hipError_t launch_sweep_x(int cpl, const Grid &g, const Compact &cv, const double *Tn, double *Ts,
                          const uint8_t *vmap, const double *vtab, const double *sfield,
                          hipStream_t s) {
  switch (cpl) {
    case 1: return launch_x_t<1>(g, cv, Tn, Ts, vmap, vtab, sfield, s);
    case 2: return launch_x_t<2>(g, cv, Tn, Ts, vmap, vtab, sfield, s);
    case 4: return launch_x_t<4>(g, cv, Tn, Ts, vmap, vtab, sfield, s);
    case 8: return launch_x_t<8>(g, cv, Tn, Ts, vmap, vtab, sfield, s);
    case 16: return launch_x_t<16>(g, cv, Tn, Ts, vmap, vtab, sfield, s);
    default: return hipErrorInvalidValue;
  }
}

size_t strided_lds_bytes(const Compact &cv, int P, int CW) {
  const int tabpad = (cv.tab_len + 1) & ~1;
  return sizeof(double) * ((size_t)tabpad + 3 * 2 * (size_t)P * CW);
}

template <int SEG>
static hipError_t launch_s_t(const Grid &g, const Compact &cv, int axis, int CW, const double *Tn,
                             const double *Y, double *X, double *surf, double dz, hipStream_t s) {
  const int L = axis == 1 ? g.ny : g.nz;
  const int P = (L + SEG - 1) / SEG;
  const int other = axis == 1 ? g.nz : g.ny;
  dim3 grid((g.nx + CW - 1) / CW, other);
  hipLaunchKernelGGL(k_sweep_strided<SEG>, grid, dim3(P * CW), strided_lds_bytes(cv, P, CW), s, g,
                     cv, axis, CW, Tn, Y, X, surf, dz);
  return hipGetLastError();
}

hipError_t launch_sweep_strided(int seg, int cw, const Grid &g, const Compact &cv, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s) {
  switch (seg) {
    case 2: return launch_s_t<2>(g, cv, axis, cw, Tn, Y, X, surf, dz, s);
    case 4: return launch_s_t<4>(g, cv, axis, cw, Tn, Y, X, surf, dz, s);
    case 8: return launch_s_t<8>(g, cv, axis, cw, Tn, Y, X, surf, dz, s);
    case 16: return launch_s_t<16>(g, cv, axis, cw, Tn, Y, X, surf, dz, s);
    case 32: return launch_s_t<32>(g, cv, axis, cw, Tn, Y, X, surf, dz, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_expand(const Grid &g, const Compact &cv, double *gz, double *gy, double *gx,
                         double *M, uint8_t *cflag, hipStream_t s) {
  const int n = g.nz * g.ny * g.nx;
  int blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_expand, dim3(blocks), dim3(256), 0, s, g, cv, gz, gy, gx, M, cflag);
  return hipGetLastError();
}

hipError_t launch_surface(const double *T, int ny, int nx, double dz, double *out, hipStream_t s) {
  const int sxy = nx * ny;
  int blocks = (sxy + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(256), 0, s, T, ny, nx, dz, out);
  return hipGetLastError();
}

}  // namespace hs2
