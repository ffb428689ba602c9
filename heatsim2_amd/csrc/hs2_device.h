// Device-side building blocks shared by the heatsim2_amd kernels (gfx950).
//
// Coefficient model.  For a COMPUTE cell c the explicit operator on axis a is
//   (L_a u)_c = g_a(c) (u_{c-e_a} - u_c) + g_a(c+e_a) (u_{c+e_a} - u_c)
// where g_a(c) is the conductance per volume (W m^-3 K^-1) of the face
// between c-e_a and c (0 on the outer faces), and M_c = rho*c/dt.  A FIXED
// cell has M = 1 and its own row has no couplings (row weight w = 0; w = 1
// for COMPUTE).  g is a property of the face: the same number from both
// COMPUTE sides (crank_nicolson.pyx:298-386 derives it per cell).
//
// Each cell therefore carries, per axis, an entry (M, w, g_minus).  Two
// views deliver it:
//   ClassView: one uint8 class id per cell per axis (built once on the
//              device from the upstream uint8 maps) indexing a <=256-entry
//              table held in LDS -- 1 byte per cell per sweep;
//   DenseView: per-cell float64 arrays M, g_z, g_y, g_x and a uint8 w
//              (per-voxel parameters, e.g. for gradients).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace hs2 {

constexpr int kWave = 64;
constexpr int kMaxWaves = 1024 / kWave;  // waves in the largest block
constexpr int kMaxVtab = 256;
constexpr int kMaxClass = 256;
constexpr int kTabStride = 4;  // doubles per class-table entry: M, w, g, pad

struct Grid {
  int nz, ny, nx;
};

struct Entry {
  double M, w, g;
  double gp;  // plus-face conductance as the cell sees it (views with kAsym only)
};

// ---------------------------------------------------------------- views
struct ClassView {
  static constexpr bool kAsym = false;  // a cell's plus face is its neighbour's minus face
  const uint8_t *__restrict__ cls[3];  // per axis 0=z, 1=y, 2=x; n each
  const double *__restrict__ tab;      // global tables, axis a at tab + off[a]
  int off[4];                          // doubles; off[3] = total length
  using Raw = uint32_t;
  // doubles of LDS for the tables of all axes (axis < 0) or one axis
  __host__ __device__ int lds_doubles(int axis) const {
    return axis < 0 ? off[3] : off[axis + 1] - off[axis];
  }
  // copy all tables (axis < 0) or one axis's table to lds
  __device__ void load_tables(double *lds, int axis) const {
    const int a0 = axis < 0 ? 0 : off[axis];
    const int len = lds_doubles(axis);
    for (int t = threadIdx.x; t < len; t += blockDim.x) lds[t] = tab[a0 + t];
  }
  // table of `axis` inside lds filled by load_tables(lds, -1)
  __device__ __forceinline__ const double *table(const double *lds, int axis) const {
    return lds + off[axis];
  }
  __device__ __forceinline__ Raw fetch(int axis, int c) const { return cls[axis][c]; }
  // four consecutive cells starting at c (c % 4 == 0)
  __device__ __forceinline__ void fetch4(int axis, int c, Raw (&r)[4]) const {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(cls[axis] + c);
    r[0] = w & 0xffu;
    r[1] = (w >> 8) & 0xffu;
    r[2] = (w >> 16) & 0xffu;
    r[3] = w >> 24;
  }
  __device__ __forceinline__ Entry resolve(const double *t, Raw r) const {
    const double2 mw = *reinterpret_cast<const double2 *>(t + r * kTabStride);
    return Entry{mw.x, mw.y, t[r * kTabStride + 2]};
  }
};

// One uint8 id per cell for all three axes: the id of the cell's (z, y, x)
// class triple, indexing a <=256-entry table whose entry holds the three
// axis entries (M, w, g, pad) side by side, x at 0, y at 4, z at 8.  For the
// x sweep, which needs every axis of a cell: one class plane instead of
// three (built by the plan when the triples number 256 or fewer).
struct CombView {
  static constexpr bool kAsym = false;
  static constexpr int kStride = 12;  // doubles per id
  const uint8_t *__restrict__ comb;
  const double *__restrict__ tab;
  int n;  // ids
  using Raw = uint32_t;
  __host__ __device__ int lds_doubles(int) const { return kStride * n; }
  __device__ void load_tables(double *lds, int) const {
    for (int t = threadIdx.x; t < kStride * n; t += blockDim.x) lds[t] = tab[t];
  }
  __device__ __forceinline__ const double *table(const double *lds, int axis) const {
    return lds + 4 * (2 - axis);
  }
  __device__ __forceinline__ Raw fetch(int, int c) const { return comb[c]; }
  __device__ __forceinline__ void fetch4(int, int c, Raw (&r)[4]) const {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(comb + c);
    r[0] = w & 0xffu;
    r[1] = (w >> 8) & 0xffu;
    r[2] = (w >> 16) & 0xffu;
    r[3] = w >> 24;
  }
  __device__ __forceinline__ Entry resolve(const double *t, Raw r) const {
    const double2 mw = *reinterpret_cast<const double2 *>(t + r * kStride);
    return Entry{mw.x, mw.y, t[r * kStride + 2]};
  }
};

struct DenseRaw {
  double M, g;
  uint32_t w;
};

struct DenseView {
  static constexpr bool kAsym = false;
  const double *__restrict__ M;
  const uint8_t *__restrict__ w;
  const double *__restrict__ g[3];
  using Raw = DenseRaw;
  __host__ __device__ int lds_doubles(int) const { return 0; }
  __device__ void load_tables(double *, int) const {}
  __device__ __forceinline__ const double *table(const double *lds, int) const { return lds; }
  __device__ __forceinline__ Raw fetch(int axis, int c) const {
    return DenseRaw{M[c], g[axis][c], w[c]};
  }
  __device__ __forceinline__ void fetch4(int axis, int c, Raw (&r)[4]) const {
    const double2 m01 = *reinterpret_cast<const double2 *>(M + c);
    const double2 m23 = *reinterpret_cast<const double2 *>(M + c + 2);
    const double2 g01 = *reinterpret_cast<const double2 *>(g[axis] + c);
    const double2 g23 = *reinterpret_cast<const double2 *>(g[axis] + c + 2);
    const uint32_t ww = *reinterpret_cast<const uint32_t *>(w + c);
    r[0] = DenseRaw{m01.x, g01.x, ww & 0xffu};
    r[1] = DenseRaw{m01.y, g01.y, (ww >> 8) & 0xffu};
    r[2] = DenseRaw{m23.x, g23.x, (ww >> 16) & 0xffu};
    r[3] = DenseRaw{m23.y, g23.y, ww >> 24};
  }
  __device__ __forceinline__ Entry resolve(const double *, Raw r) const {
    return Entry{r.M, (double)r.w, r.g};
  }
};

struct AsymRaw {
  double M, g, gp;
  uint32_t w;
};

// Per-voxel coefficients where the two cells of a face see it differently
// (curvature mode: each cell weighs a face by its own area and volume), so
// each cell carries its plus-face conductance besides its minus-face one.
struct AsymView {
  static constexpr bool kAsym = true;
  const double *__restrict__ M;
  const uint8_t *__restrict__ w;
  const double *__restrict__ g[3];   // minus faces
  const double *__restrict__ gp[3];  // plus faces
  using Raw = AsymRaw;
  __host__ __device__ int lds_doubles(int) const { return 0; }
  __device__ void load_tables(double *, int) const {}
  __device__ __forceinline__ const double *table(const double *lds, int) const { return lds; }
  __device__ __forceinline__ Raw fetch(int axis, int c) const {
    return AsymRaw{M[c], g[axis][c], gp[axis][c], w[c]};
  }
  __device__ __forceinline__ void fetch4(int axis, int c, Raw (&r)[4]) const {
#pragma unroll
    for (int q = 0; q < 4; q++) r[q] = fetch(axis, c + q);
  }
  __device__ __forceinline__ Entry resolve(const double *, Raw r) const {
    return Entry{r.M, (double)r.w, r.g, r.gp};
  }
};

// ------------------------------------------------------ buffer access
// Buffer descriptors (gfx9 family: third dword 0x00020000) give loads and
// stores a 32-bit byte offset from a scalar base; accesses beyond
// num_records read 0 and are dropped.
using BufRsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ BufRsrc make_buf(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ double2 buf_load_d2(BufRsrc r, unsigned off) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ double buf_load_d(BufRsrc r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t buf_load_u32(BufRsrc r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ uint32_t buf_load_u8(BufRsrc r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
__device__ __forceinline__ void buf_store_d2(double2 v, BufRsrc r, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b128(
      __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, off, 0, 0);
}
// the four bytes of a class word, lowest first
__device__ __forceinline__ void unpack4(uint32_t w, uint32_t (&r)[4]) {
  r[0] = w & 0xffu;
  r[1] = (w >> 8) & 0xffu;
  r[2] = (w >> 16) & 0xffu;
  r[3] = w >> 24;
}

// Logical block index that keeps consecutive logical blocks on one XCD.
// Hardware hands block b to XCD b % 8; give each XCD a contiguous range.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int x = b & 7, idx = b >> 3;
  return x * q + (x < r ? x : r) + idx;
}

// The same in chunks: consecutive logical blocks go to one XCD C at a time,
// and the chunks go round the XCDs (blocks past the last whole round of 8 C
// keep their index).  A bijection on [0, nblocks).
template <int C>
__device__ __forceinline__ int xcd_remap_chunks(int b, int nblocks) {
  constexpr int R = 8 * C;
  if (b >= nblocks / R * R) return b;
  const int x = b & 7, k = b >> 3;
  return ((k / C) * 8 + x) * C + k % C;
}

// ------------------------------------------------------- lane exchange
// Within groups of G lanes (G divides 64).  up(v, s): value of lane l-s of
// the same group, 0 for the first s lanes; down(v, s): lane l+s, 0 for the
// last s lanes.  G = 16 uses DPP row shifts (a VALU move, no LDS traffic);
// wider groups use ds_bpermute.
template <int G>
struct LaneX {
  static __device__ __forceinline__ double up(double v, int s, int lig);
  static __device__ __forceinline__ double down(double v, int s, int lig);
};

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// row_shr:s = 0x110+s (lane l reads l-s), row_shl:s = 0x100+s (reads l+s)
template <>
struct LaneX<16> {
  static __device__ __forceinline__ double up(double v, int s, int) {
    switch (s) {
      case 1: return dpp_d<0x111>(v);
      case 2: return dpp_d<0x112>(v);
      case 4: return dpp_d<0x114>(v);
      default: return dpp_d<0x118>(v);
    }
  }
  static __device__ __forceinline__ double down(double v, int s, int) {
    switch (s) {
      case 1: return dpp_d<0x101>(v);
      case 2: return dpp_d<0x102>(v);
      case 4: return dpp_d<0x104>(v);
      default: return dpp_d<0x108>(v);
    }
  }
};

template <int G>
__device__ __forceinline__ double LaneX<G>::up(double v, int s, int lig) {
  const double x = __shfl_up(v, (unsigned)s, kWave);
  return lig >= s ? x : 0.0;
}
template <int G>
__device__ __forceinline__ double LaneX<G>::down(double v, int s, int lig) {
  const double x = __shfl_down(v, (unsigned)s, kWave);
  return lig + s < G ? x : 0.0;
}

// 1/x from the hardware reciprocal refined by two Newton steps.  Measured on
// gfx950 over 2^20 values in [2^-30, 2^30]: v_rcp_f64 alone is good to
// ~2^-24.4 relative, one step leaves up to ~20 x 2^-53, two steps 1 x 2^-53
// (as IEEE division).  The one-step error accumulated visibly over a
// thousand time steps.  x is a well-scaled positive pivot here, never 0,
// inf or denormal.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Couplings below this magnitude cannot change an O(1)-relative solution by
// even one ulp (2^-64 << 2^-53), so the reduced system is already diagonal.
constexpr double kNegligible = 5.421010862427522e-20;  // 2^-64

// One PCR reduction of a normalised row  a x_{i-1} + x_i + c x_{i+1} = r
// with its neighbour rows L (i-1) and R (i+1), giving a row that couples
// i-2 and i+2.  Absent neighbours are passed as zeros.
__device__ __forceinline__ void pcr_combine(double &a, double &c, double &r, double aL,
                                            double cL, double rL, double aR, double cR,
                                            double rR) {
  const double k = rcp(1.0 - a * cL - c * aR);
  const double an = -a * aL * k;
  const double cn = -c * cR * k;
  const double rn = (r - a * rL - c * rR) * k;
  a = an;
  c = cn;
  r = rn;
}

// Segment elimination of rows already normalised to unit diagonal,
// a_q x_{q-1} + x_q + c_q x_{q+1} = r_q, q = 0..N-1 of one contiguous segment.
// Leaves the two end rows
//   (As, Cs, Rs): As x_{before} + x_0     + Cs x_{N-1}  = Rs
//   (Ae, Ce, Re): Ae x_0        + x_{N-1} + Ce x_{after} = Re
// and, in place for interior rows 1..N-2, a[q] x_0 + x_q + c[q] x_{N-1} = r[q].
// Forward pass eliminates x_{q-1} from rows 2..N-1 (expressing them in x_0),
// backward pass eliminates x_{q+1} from rows N-3..1 (in x_{N-1}).  N >= 2.
template <int N>
__device__ __forceinline__ void segment_reduce(double (&a)[N], double (&c)[N], double (&r)[N],
                                               double &As, double &Cs, double &Rs, double &Ae,
                                               double &Ce, double &Re) {
  static_assert(N >= 2, "segment needs two rows");
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

// segment_reduce on rows O..O+M-1 of the arrays, in place (same algebra,
// indices shifted): leaves the end rows of that sub-segment and its interior
// rows in terms of x_O and x_{O+M-1}.
template <int N, int O, int M>
__device__ __forceinline__ void sub_reduce(double (&a)[N], double (&c)[N], double (&r)[N],
                                           double &As, double &Cs, double &Rs, double &Ae,
                                           double &Ce, double &Re) {
  static_assert(M >= 3 && O + M <= N, "sub-segment of three rows or more");
#pragma unroll
  for (int q = O + 2; q < O + M; q++) {
    const double k = 1.0 / (1.0 - a[q] * c[q - 1]);
    const double an = -a[q] * a[q - 1] * k;
    c[q] = c[q] * k;
    r[q] = (r[q] - a[q] * r[q - 1]) * k;
    a[q] = an;
  }
  Ae = a[O + M - 1];
  Ce = c[O + M - 1];
  Re = r[O + M - 1];
#pragma unroll
  for (int q = O + M - 3; q >= O + 1; q--) {
    a[q] = a[q] - c[q] * a[q + 1];
    r[q] = r[q] - c[q] * r[q + 1];
    c[q] = -c[q] * c[q + 1];
  }
  const double k = 1.0 / (1.0 - c[O] * a[O + 1]);
  As = a[O] * k;
  Cs = -c[O] * c[O + 1] * k;
  Rs = (r[O] - c[O] * r[O + 1]) * k;
}

// The two inner end cells of a segment reduced in halves, as affine
// functions of the segment's end cells: x_{h-1} = P1 + A1 x_0 + C1 x_{N-1},
// x_h = P2 + A2 x_0 + C2 x_{N-1} (h = N/2).
struct HalfJoin {
  double P1, A1, C1, P2, A2, C2;
};

// segment_reduce by halves: rows 0..h-1 and h..N-1 are reduced
// independently (two dependency chains instead of one), then joined.  The
// join eliminates x_{h-1} and x_h between the halves' four end rows
//   As1 x_b + x_0 + Cs1 x_{h-1} = Rs1,   Ae1 x_0 + x_{h-1} + Ce1 x_h = Re1,
//   As2 x_{h-1} + x_h + Cs2 x_{N-1} = Rs2,   Ae2 x_h + x_{N-1} + Ce2 x_a = Re2,
// leaving the same two end rows as segment_reduce; interior rows stay in
// terms of their half's ends (x_0, x_{h-1}) and (x_h, x_{N-1}).
template <int N>
__device__ __forceinline__ void segment_reduce_halves(double (&a)[N], double (&c)[N],
                                                      double (&r)[N], double &As, double &Cs,
                                                      double &Rs, double &Ae, double &Ce,
                                                      double &Re, HalfJoin &j) {
  static_assert(N % 2 == 0 && N >= 6, "two halves of three rows or more");
  constexpr int H = N / 2;
  double As1, Cs1, Rs1, Ae1, Ce1, Re1, As2, Cs2, Rs2, Ae2, Ce2, Re2;
  sub_reduce<N, 0, H>(a, c, r, As1, Cs1, Rs1, Ae1, Ce1, Re1);
  sub_reduce<N, H, H>(a, c, r, As2, Cs2, Rs2, Ae2, Ce2, Re2);
  // x_{h-1} + Ce1 x_h = Re1 - Ae1 x_0 and As2 x_{h-1} + x_h = Rs2 - Cs2 x_{N-1}
  const double k = 1.0 / (1.0 - Ce1 * As2);
  j.P1 = (Re1 - Ce1 * Rs2) * k;
  j.A1 = -Ae1 * k;
  j.C1 = Ce1 * Cs2 * k;
  j.P2 = (Rs2 - As2 * Re1) * k;
  j.A2 = As2 * Ae1 * k;
  j.C2 = -Cs2 * k;
  const double d0 = 1.0 / (1.0 + Cs1 * j.A1);
  As = As1 * d0;
  Cs = Cs1 * j.C1 * d0;
  Rs = (Rs1 - Cs1 * j.P1) * d0;
  const double d1 = 1.0 / (1.0 + Ae2 * j.C2);
  Ae = Ae2 * j.A2 * d1;
  Ce = Ce2 * d1;
  Re = (Re2 - Ae2 * j.P2) * d1;
}

// Segment elimination, streamed row by row.  Rows q = 0..N-1 of one
// contiguous segment, a_q x_{q-1} + b_q x_q + c_q x_{q+1} = r_q (b_q > 0),
// are fed in order with row(q, ...); one reciprocal per row.  finish() then
// leaves
//   (As, Cs, Rs): As x_{before} + x_0     + Cs x_{N-1}  = Rs
//   (Ae, Ce, Re): Ae x_0        + x_{N-1} + Ce x_{after} = Re
// and, for interior rows 1..N-2, a[q] x_0 + x_q + c[q] x_{N-1} = r[q].
// N >= 2.
template <int N>
struct Seg {
  static_assert(N >= 2, "segment needs two rows");
  double a[N], c[N], r[N];
  double b0;
  __device__ __forceinline__ void row(int q, double aq, double bq, double cq, double rq) {
    if (q == 0) {
      a[0] = aq;
      b0 = bq;
      c[0] = cq;
      r[0] = rq;
    } else if (q == 1) {
      const double k = rcp(bq);
      a[1] = aq * k;
      c[1] = cq * k;
      r[1] = rq * k;
    } else {
      const double k = rcp(bq - aq * c[q - 1]);
      a[q] = -aq * a[q - 1] * k;
      c[q] = cq * k;
      r[q] = (rq - aq * r[q - 1]) * k;
    }
  }
  __device__ __forceinline__ void finish(double &As, double &Cs, double &Rs, double &Ae,
                                         double &Ce, double &Re) {
    Ae = a[N - 1];
    Ce = c[N - 1];
    Re = r[N - 1];
    if constexpr (N == 2) {
      const double k = rcp(b0);
      As = a[0] * k;
      Cs = c[0] * k;
      Rs = r[0] * k;
    } else {
#pragma unroll
      for (int q = N - 3; q >= 1; q--) {
        a[q] = a[q] - c[q] * a[q + 1];
        r[q] = r[q] - c[q] * r[q + 1];
        c[q] = -c[q] * c[q + 1];
      }
      const double k = rcp(b0 - c[0] * a[1]);
      As = a[0] * k;
      Cs = -c[0] * c[1] * k;
      Rs = (r[0] - c[0] * r[1]) * k;
    }
  }
  // interior solution once the segment's end values are known
  __device__ __forceinline__ double interior(int q, double xs, double xe) const {
    return r[q] - a[q] * xs - c[q] * xe;
  }
};

__device__ __forceinline__ double coupling(double As, double Cs, double Ae, double Ce) {
  return fmax(fmax(fabs(As), fabs(Cs)), fmax(fabs(Ae), fabs(Ce)));
}

// PCR across the threads of a block over the 2P end rows of P segments per
// line, CW lines side by side in LDS (row r of line cl at r*CW + cl); thread
// (p, cl) owns rows 2p and 2p+1, written, with a barrier passed, on entry.
// On return rr holds the end rows' solutions.  After the first level every
// coupling is a near one times one across a whole segment; once all of them
// are negligible (usual on a diagonally dominant line) the other levels are
// skipped.  The block-wide decision rides on the level's own barrier: each
// wave posts its ballot before it, every thread reads all posts after it.
__device__ __forceinline__ void lds_pcr_ends(double *ra, double *rc, double *rr, int p, int P,
                                             int CW, int cl, bool act = true) {
  // act: this thread holds segment p < P (threads past the block's P x CW
  // segments only take part in the barriers)
  __shared__ int wave_more[kMaxWaves];
  const int R = 2 * P;
  for (int d = 1; d < R; d <<= 1) {
    double na[2], nc[2], nr[2];
    if (act) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = 2 * p + h;
        double aa = ra[row * CW + cl], cc = rc[row * CW + cl], r0 = rr[row * CW + cl];
        const int rl = row - d >= 0 ? row - d : row, rh = row + d < R ? row + d : row;
        const double m0 = row - d >= 0 ? 1.0 : 0.0, m1 = row + d < R ? 1.0 : 0.0;
        const double aL = m0 * ra[rl * CW + cl], cL = m0 * rc[rl * CW + cl], rL = m0 * rr[rl * CW + cl];
        const double aR = m1 * ra[rh * CW + cl], cR = m1 * rc[rh * CW + cl], rR = m1 * rr[rh * CW + cl];
        pcr_combine(aa, cc, r0, aL, cL, rL, aR, cR, rR);
        na[h] = aa; nc[h] = cc; nr[h] = r0;
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = 2 * p + h;
        ra[row * CW + cl] = na[h]; rc[row * CW + cl] = nc[h]; rr[row * CW + cl] = nr[h];
      }
    }
    if (d == 1) {
      const bool more = act && coupling(na[0], nc[0], na[1], nc[1]) > kNegligible;
      const unsigned long long bal = __ballot(more);
      if ((threadIdx.x & (kWave - 1)) == 0) wave_more[threadIdx.x / kWave] = bal != 0ull;
    }
    __syncthreads();
    if (d == 1) {
      int any = 0;
      for (int w = 0; w < (int)(blockDim.x + kWave - 1) / kWave; w++) any |= wave_more[w];
      if (!any) break;
    }
  }
}

// Single-row "segment" (CPL = 1): the row itself, normalised.
struct Row1 {
  double a, c, r;
  __device__ __forceinline__ void row(int, double aq, double bq, double cq, double rq) {
    const double k = rcp(bq);
    a = aq * k;
    c = cq * k;
    r = rq * k;
  }
};

template <int CPL>
using SegOf = typename std::conditional<CPL == 1, Row1, Seg<CPL>>::type;

// Solve the G-lane group's tridiagonal system: each lane holds CPL
// consecutive unknowns whose rows were fed to `s`; x receives the solution.
// The cross-lane reduction is skipped when every coupling left after the
// in-lane elimination is negligible for the whole wavefront (a
// wave-uniform decision).
template <int CPL, int G>
__device__ __forceinline__ void group_solve(SegOf<CPL> &s, double (&x)[CPL], int lig) {
  using X = LaneX<G>;
  if constexpr (CPL == 1) {
    double aa = s.a, cc = s.c, rr = s.r;
    if (__any(fmax(fabs(aa), fabs(cc)) > kNegligible)) {
#pragma unroll
      for (int d = 1; d < G; d <<= 1) {
        const double aL = X::up(aa, d, lig), cL = X::up(cc, d, lig), rL = X::up(rr, d, lig);
        const double aR = X::down(aa, d, lig), cR = X::down(cc, d, lig), rR = X::down(rr, d, lig);
        pcr_combine(aa, cc, rr, aL, cL, rL, aR, cR, rR);
      }
    }
    x[0] = rr;
  } else {
    double As, Cs, Rs, Ae, Ce, Re;
    s.finish(As, Cs, Rs, Ae, Ce, Re);
    if (__any(coupling(As, Cs, Ae, Ce) > kNegligible)) {
      // reduced order s0 e0 s1 e1 ...: the first PCR step pairs s_p with
      // e_{p-1}, e_p and e_p with s_p, s_{p+1}, splitting it into an s- and
      // an e-system that are then reduced independently across the lanes.
      {
        const double aL = X::up(Ae, 1, lig), cL = X::up(Ce, 1, lig), rL = X::up(Re, 1, lig);
        const double aR = X::down(As, 1, lig), cR = X::down(Cs, 1, lig), rR = X::down(Rs, 1, lig);
        double as = As, cs = Cs, rs = Rs, ae = Ae, ce = Ce, re = Re;
        pcr_combine(as, cs, rs, aL, cL, rL, Ae, Ce, Re);
        pcr_combine(ae, ce, re, As, Cs, Rs, aR, cR, rR);
        As = as; Cs = cs; Rs = rs; Ae = ae; Ce = ce; Re = re;
      }
      // after the pairing each coupling is a near one times one across a
      // whole segment: on a diagonally dominant line these are usually
      // negligible already, and the remaining levels are skipped
      if (__any(coupling(As, Cs, Ae, Ce) > kNegligible)) {
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
          const double aL = X::up(As, d, lig), cL = X::up(Cs, d, lig), rL = X::up(Rs, d, lig);
          const double aR = X::down(As, d, lig), cR = X::down(Cs, d, lig), rR = X::down(Rs, d, lig);
          const double eaL = X::up(Ae, d, lig), ecL = X::up(Ce, d, lig), erL = X::up(Re, d, lig);
          const double eaR = X::down(Ae, d, lig), ecR = X::down(Ce, d, lig), erR = X::down(Re, d, lig);
          pcr_combine(As, Cs, Rs, aL, cL, rL, aR, cR, rR);
          pcr_combine(Ae, Ce, Re, eaL, ecL, erL, eaR, ecR, erR);
        }
      }
    }
    x[0] = Rs;
    x[CPL - 1] = Re;
#pragma unroll
    for (int q = 1; q < CPL - 1; q++) x[q] = s.interior(q, Rs, Re);
  }
}

// Seg with the right-hand sides kept in the lane's own LDS slots rp[0..N-1]
// instead of registers (the same elimination; the solution is written back
// into the slots).  Frees 2N VGPRs in kernels whose occupancy the
// registers limit.
template <int N>
struct SegLds {
  static_assert(N >= 2, "segment needs two rows");
  double a[N], c[N];
  double b0, rprev;
  double *rp;
  __device__ __forceinline__ void row(int q, double aq, double bq, double cq, double rq) {
    if (q == 0) {
      a[0] = aq;
      b0 = bq;
      c[0] = cq;
      rprev = rq;
    } else if (q == 1) {
      const double k = rcp(bq);
      a[1] = aq * k;
      c[1] = cq * k;
      rprev = rq * k;
    } else {
      const double k = rcp(bq - aq * c[q - 1]);
      a[q] = -aq * a[q - 1] * k;
      c[q] = cq * k;
      rprev = (rq - aq * rprev) * k;
    }
    rp[q] = rprev;
  }
  __device__ __forceinline__ void finish(double &As, double &Cs, double &Rs, double &Ae,
                                         double &Ce, double &Re) {
    Ae = a[N - 1];
    Ce = c[N - 1];
    Re = rprev;
    if constexpr (N == 2) {
      const double k = rcp(b0);
      As = a[0] * k;
      Cs = c[0] * k;
      Rs = rp[0] * k;
    } else {
      double rnext = rp[N - 2];
#pragma unroll
      for (int q = N - 3; q >= 1; q--) {
        a[q] = a[q] - c[q] * a[q + 1];
        const double rq = rp[q] - c[q] * rnext;
        rp[q] = rq;
        rnext = rq;
        c[q] = -c[q] * c[q + 1];
      }
      const double k = rcp(b0 - c[0] * a[1]);
      As = a[0] * k;
      Cs = -c[0] * c[1] * k;
      Rs = (rp[0] - c[0] * rnext) * k;
    }
  }
};

// group_solve for a SegLds: the solution overwrites the lane's slots.
template <int CPL, int G>
__device__ __forceinline__ void group_solve_lds(SegLds<CPL> &s, int lig) {
  using X = LaneX<G>;
  double As, Cs, Rs, Ae, Ce, Re;
  s.finish(As, Cs, Rs, Ae, Ce, Re);
  if (__any(coupling(As, Cs, Ae, Ce) > kNegligible)) {
    {
      const double aL = X::up(Ae, 1, lig), cL = X::up(Ce, 1, lig), rL = X::up(Re, 1, lig);
      const double aR = X::down(As, 1, lig), cR = X::down(Cs, 1, lig), rR = X::down(Rs, 1, lig);
      double as = As, cs = Cs, rs = Rs, ae = Ae, ce = Ce, re = Re;
      pcr_combine(as, cs, rs, aL, cL, rL, Ae, Ce, Re);
      pcr_combine(ae, ce, re, As, Cs, Rs, aR, cR, rR);
      As = as; Cs = cs; Rs = rs; Ae = ae; Ce = ce; Re = re;
    }
    if (__any(coupling(As, Cs, Ae, Ce) > kNegligible)) {  // as in group_solve
#pragma unroll
      for (int d = 1; d < G; d <<= 1) {
        const double aL = X::up(As, d, lig), cL = X::up(Cs, d, lig), rL = X::up(Rs, d, lig);
        const double aR = X::down(As, d, lig), cR = X::down(Cs, d, lig), rR = X::down(Rs, d, lig);
        const double eaL = X::up(Ae, d, lig), ecL = X::up(Ce, d, lig), erL = X::up(Re, d, lig);
        const double eaR = X::down(Ae, d, lig), ecR = X::down(Ce, d, lig), erR = X::down(Re, d, lig);
        pcr_combine(As, Cs, Rs, aL, cL, rL, aR, cR, rR);
        pcr_combine(Ae, Ce, Re, eaL, ecL, erL, eaR, ecR, erR);
      }
    }
  }
#pragma unroll
  for (int q = 1; q < CPL - 1; q++) s.rp[q] = s.rp[q] - s.a[q] * Rs - s.c[q] * Re;
  s.rp[0] = Rs;
  s.rp[CPL - 1] = Re;
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
