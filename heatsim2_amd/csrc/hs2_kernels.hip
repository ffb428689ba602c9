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
// x (contiguous): a group of G lanes per x-line, CPL consecutive cells per
//   lane; each lane eliminates its cells locally and the 2G-row reduced
//   system is solved by parallel cyclic reduction across the group (DPP row
//   shifts for G = 16).
// y, z (strided): CW consecutive lines per block, coalesced along x; each
//   thread runs the Thomas-style elimination of one SEG-cell segment of one
//   line in registers, the 2P-row reduced system of each line is solved by
//   PCR in LDS, then each thread back-substitutes its segment.  With one
//   segment per line this is the plain one-thread-per-line Thomas sweep.
//
// Every global load goes through a clamped in-range address with no branch
// around it, so a thread issues all of its loads before it first waits.
#include <algorithm>

#include "hs2_device.h"
#include "hs2_batch.h"
#include "hs2_launch.h"

namespace hs2 {

// Conductance of a cell's plus face on one axis: the cell's own (views whose
// cells see a face differently from either side) or else the minus face of
// the next cell, which is the same face.
template <class V>
__device__ __forceinline__ double plus_g(const V &view, const double *tab, typename V::Raw own,
                                         typename V::Raw next) {
  if constexpr (V::kAsym) return view.resolve(tab, own).gp;
  else return view.resolve(tab, next).g;
}

// ---------------------------------------------------------------- x sweep
template <class V, int CPL, int G>
__global__ __launch_bounds__(256) void k_sweep_x(Grid g, V view, const double *__restrict__ Tn,
                                                 double *__restrict__ Ts,
                                                 const uint8_t *__restrict__ vmap,
                                                 const double *__restrict__ vtab_g,
                                                 const double *__restrict__ sfield) {
  extern __shared__ double lds[];
  view.load_tables(lds, -1);
  double *vtab = lds + ((view.lds_doubles(-1) + 1) & ~1);
  if (vmap)
    for (int t = threadIdx.x; t < kMaxVtab; t += blockDim.x) vtab[t] = vtab_g[t];
  __syncthreads();
  const double *tabz = view.table(lds, 0), *taby = view.table(lds, 1), *tabx = view.table(lds, 2);

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nlines = nz * ny;
  constexpr int LPB = 256 / G;  // lines per block
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int lig = threadIdx.x % G;
  const int line = b * LPB + threadIdx.x / G;
  const bool lvalid = line < nlines;
  const int lc = lvalid ? line : nlines - 1;
  const int k = lc / ny, j = lc - k * ny;
  const int base = lc * nx;
  const int bym = (j > 0 ? lc - 1 : lc) * nx, byp = (j < ny - 1 ? lc + 1 : lc) * nx;
  const int bzm = (k > 0 ? lc - ny : lc) * nx, bzp = (k < nz - 1 ? lc + ny : lc) * nx;
  const int i0 = lig * CPL;

  // ---- loads (all issued before first use)
  double T[CPL], Tym[CPL], Typ[CPL], Tzm[CPL], Tzp[CPL];
  typename V::Raw rx[CPL], ry[CPL], ryp[CPL], rz[CPL], rzp[CPL];
#pragma unroll
  for (int q = 0; q < CPL; q++) {
    const int i = min(i0 + q, nx - 1);
    T[q] = Tn[base + i];
    Tym[q] = Tn[bym + i];
    Typ[q] = Tn[byp + i];
    Tzm[q] = Tn[bzm + i];
    Tzp[q] = Tn[bzp + i];
    rx[q] = view.fetch(2, base + i);
    ry[q] = view.fetch(1, base + i);
    ryp[q] = V::kAsym ? ry[q] : view.fetch(1, byp + i);
    rz[q] = view.fetch(0, base + i);
    rzp[q] = V::kAsym ? rz[q] : view.fetch(0, bzp + i);
  }
  double S[CPL];
#pragma unroll
  for (int q = 0; q < CPL; q++) S[q] = 0.0;
  if (vmap) {
#pragma unroll
    for (int q = 0; q < CPL; q++) S[q] += vtab[vmap[base + min(i0 + q, nx - 1)]];
  }
  if (sfield) {
#pragma unroll
    for (int q = 0; q < CPL; q++) S[q] += sfield[base + min(i0 + q, nx - 1)];
  }

  // ---- coefficients and right-hand side
  Entry ex[CPL + 1];
#pragma unroll
  for (int q = 0; q < CPL; q++) ex[q] = view.resolve(tabx, rx[q]);
  // the face after the lane's last cell belongs to the next lane's first cell
  ex[CPL].g = LaneX<G>::down(ex[0].g, 1, lig);
  const double TL = LaneX<G>::up(T[CPL - 1], 1, lig);
  const double TR = LaneX<G>::down(T[0], 1, lig);
  SegOf<CPL> seg;
#pragma unroll
  for (int q = 0; q < CPL; q++) {
    const int i = i0 + q;
    const bool valid = lvalid && i < nx;
    const double M = ex[q].M, w = ex[q].w;
    const double gxm = i > 0 ? ex[q].g : 0.0;
    const double gxp = i < nx - 1 ? (V::kAsym ? ex[q].gp : ex[q + 1].g) : 0.0;
    const double gym = j > 0 ? view.resolve(taby, ry[q]).g : 0.0;
    const double gyp = j < ny - 1 ? plus_g(view, taby, ry[q], ryp[q]) : 0.0;
    const double gzm = k > 0 ? view.resolve(tabz, rz[q]).g : 0.0;
    const double gzp = k < nz - 1 ? plus_g(view, tabz, rz[q], rzp[q]) : 0.0;
    const double Tc = T[q];
    const double Txm = q > 0 ? T[q - 1] : TL;
    const double Txp = q < CPL - 1 ? T[q + 1] : TR;
    const double expl = 0.5 * (gxm * (Txm - Tc) + gxp * (Txp - Tc)) + gym * (Tym[q] - Tc) +
                        gyp * (Typ[q] - Tc) + gzm * (Tzm[q] - Tc) + gzp * (Tzp[q] - Tc) + S[q];
    const double rhs = M * Tc + w * expl;
    const double am = -0.5 * w * gxm, cp = -0.5 * w * gxp;
    seg.row(q, valid ? am : 0.0, valid ? M - am - cp : 1.0, valid ? cp : 0.0, valid ? rhs : 0.0);
  }
  double x[CPL];
  group_solve<CPL, G>(seg, x, lig);
  if (lvalid) {
#pragma unroll
    for (int q = 0; q < CPL; q++)
      if (i0 + q < nx) Ts[base + i0 + q] = x[q];
  }
}

// x sweep, two-phase form.  Phase 1 forms each cell's right-hand side with
// lanes running along x (coalesced loads of T^n and its four off-axis
// neighbour rows, U cells per thread per batch, all loads of a batch issued
// before any is used) and parks it in LDS; phase 2 reads it back as G lanes x
// SEG consecutive cells per line, recomputes the x-axis coefficients from
// the class bytes and solves like k_sweep_x.  LPB lines per block.  SRC adds
// the volumetric source terms (only on the rare steps that carry one).
// RY writes y's right-hand side instead of T*; PZ orders the block's lines
// so that a thread's quads share their y and z neighbour rows in registers.
template <class V, int SEG, int G, int LPB, bool SRC, bool FULL = false, bool BUF = false,
          bool RY = false, bool PZ = false>
__device__ __forceinline__ void sweep_x2_body(Grid g, V view, const double *__restrict__ Tn,
                                              double *__restrict__ Ts,
                                              const uint8_t *__restrict__ vmap,
                                              const double *__restrict__ vtab_g,
                                              const double *__restrict__ sfield, int blk,
                                              int ty) {
  constexpr int NT = G * LPB;
  constexpr int RS = G * SEG + ((G * SEG) >> 4) + 1;  // padded LDS row (doubles)
  extern __shared__ double lds[];
  // own, y-1, y+1, z-1, z+1 row bases.  A neighbour row outside the grid is
  // replaced by the line itself, and a missing x neighbour by the cell
  // itself: that face's difference T_nb - T_c is then exactly 0, so its
  // (finite) conductance needs no mask.  lok: the line exists (lines past
  // the grid's end load a clamped line and store nothing).
  __shared__ int lbase[LPB][5];
  __shared__ bool lok[LPB];
  view.load_tables(lds, -1);
  const int tabpad = (view.lds_doubles(-1) + 1) & ~1;
  double *vtab = lds + tabpad;
  double *rbuf = vtab + (SRC ? kMaxVtab : 0);
  // BUF: phase 1's x-class words, one per quad, for phase 2 (which would
  // otherwise fetch them again right after the barrier)
  [[maybe_unused]] uint32_t *cbuf = reinterpret_cast<uint32_t *>(rbuf + LPB * RS);
  // PARK (per-voxel coefficients): phase 1 also leaves each cell's x entry
  // (M, g, w) in LDS beside its right-hand side, so that phase 2, whose lanes
  // own runs of SEG cells, reads them there instead of from memory in
  // SEG-strided (uncoalesced) loads
  constexpr bool PARK = std::is_same<V, DenseView>::value;
  [[maybe_unused]] double *pM = rbuf + LPB * RS;
  [[maybe_unused]] double *pG = pM + LPB * RS;
  [[maybe_unused]] uint8_t *pW = reinterpret_cast<uint8_t *>(pG + LPB * RS);
  if (SRC && vmap)
    for (int t = threadIdx.x; t < kMaxVtab; t += NT) vtab[t] = vtab_g[t];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nlines = nz * ny;
  const int line0 = blk * LPB;
  const int tid = threadIdx.x;
  // BUF (class maps, field < 4 GiB): loads and stores through buffer
  // descriptors sized to the arrays
  [[maybe_unused]] BufRsrc bT, bS, bC[3];
  if constexpr (BUF) {
    const unsigned ncell = (unsigned)nlines * (unsigned)nx;
    bT = make_buf(Tn, ncell * 8u);
    bS = make_buf(Ts, ncell * 8u);
    if constexpr (std::is_same<V, CombView>::value) {
      bC[0] = bC[1] = bC[2] = make_buf(view.comb, ncell);
    } else {
#pragma unroll
      for (int a = 0; a < 3; a++) bC[a] = make_buf(view.cls[a], ncell);
    }
  }
  if (tid < LPB) {
    // ty = 0: the block's lines follow each other in (k, j) order.  ty > 0:
    // they tile LPB/ty planes x ty rows, so that lines share z neighbour rows
    // as well as y ones (fewer rows fetched per line).
    int k, j;
    bool ok;
    if (PZ) {
      // 2 W planes x 2 rows, numbered so that the two quads a thread holds
      // in one phase-1 batch (lines l and l + W, l = 2 W batch + W b + slab)
      // are rows j and j + 1 of one plane (each is the other's y neighbour),
      // and its next batch holds the same two rows one plane up
      // W = lines per quad slab (NT threads over lines of SEG*G/4 quads) =
      // LPB/4 for 16-cell lanes (4 for 16-line blocks, 1 for 4-line ones); a
      // tile is 2 W planes x 2 rows
      constexpr int W = NT / (SEG * G / 4);
      const int nbj = (ny + 1) / 2;
      const int kt = blk / nbj, jt = blk - kt * nbj;
      const int it = tid / (2 * W), b = (tid / W) % 2, wv = tid % W;
      k = kt * 2 * W + wv * 2 + it;
      j = jt * 2 + b;
      ok = k < nz && j < ny;
      k = min(k, nz - 1);
      j = min(j, ny - 1);
    } else if (ty > 0) {
      const int tz = LPB / ty, nbj = (ny + ty - 1) / ty;
      const int kt = blk / nbj, jt = blk - kt * nbj;
      k = kt * tz + tid / ty;
      j = jt * ty + tid % ty;
      ok = k < nz && j < ny;
      k = min(k, nz - 1);
      j = min(j, ny - 1);
    } else {
      const int line = line0 + tid;
      ok = line < nlines;
      const int lc = ok ? line : nlines - 1;
      k = lc / ny;
      j = lc - k * ny;
    }
    const int lc = k * ny + j;
    lok[tid] = ok;
    lbase[tid][0] = lc * nx;
    lbase[tid][1] = (j > 0 ? lc - 1 : lc) * nx;
    lbase[tid][2] = (j < ny - 1 ? lc + 1 : lc) * nx;
    lbase[tid][3] = (k > 0 ? lc - ny : lc) * nx;
    lbase[tid][4] = (k < nz - 1 ? lc + ny : lc) * nx;
  }
  __syncthreads();
  const double *tabz = view.table(lds, 0), *taby = view.table(lds, 1), *tabx = view.table(lds, 2);

  // ---- phase 1: right-hand sides, lanes along x
  // quads: 4 consecutive cells per thread, vector loads; quad e = tid + NT*u
  // of the block's LPB x nx/4 quads, walked incrementally as (l, q).  FQ
  // (full lines, nx = SEG*G): the quads per line and the trip count are
  // compile-time and every quad exists, so no bounds checks.
  constexpr bool FQ = FULL && SEG % 4 == 0;
  constexpr int UQ = FQ && SEG < 8 ? 1 : 2;
  // RY: besides the right-hand side, keep 0.5 w (Ly T^n) and the x-class word
  // of every quad until the store, which then writes y's right-hand side
  // R_y = M T* - 0.5 w (Ly T^n) in place of T*, so that the y sweep needs no
  // T^n.  The batches are unrolled to keep those in registers.
  constexpr int NIT = FQ ? SEG / (4 * UQ) : 1;
  static_assert(!RY || (FQ && BUF), "RY: full-line class-map kernel");
  static_assert(!PZ || (FQ && BUF && UQ == 2 && SEG == 16 && LPB % 4 == 0),
                "PZ: 16-cell lanes, blocks of 4, 8 or 16 lines (tiles of LPB/2 planes x 2 rows)");
  [[maybe_unused]] double hly[RY ? NIT : 1][UQ][4];
  [[maybe_unused]] uint32_t rxw[RY ? NIT : 1][UQ];
  // PZ with RY: a thread's four quads are rows j, j+1 of planes k, k+1 over
  // the two unrolled batches; the first batch's z+1 rows are the second's own
  // rows and its own rows the second's z-1 rows, so they are carried
  constexpr bool PZ4 = PZ && RY && NIT == 2;
  [[maybe_unused]] double cT1[2][4], cT2[2][4];
  [[maybe_unused]] typename V::Raw cRz2[2][4];
  // COMB: one id per cell for all three axes, so a row's ids are one word
  // (own, y+1, z+1 rows: three loads per quad instead of five), and the
  // second batch's own word is the first batch's z+1 word
  constexpr bool COMB = std::is_same<V, CombView>::value;
  [[maybe_unused]] uint32_t cWz[2];
  if ((nx & 3) == 0) {
    const int qpl = FQ ? SEG * G / 4 : nx >> 2;
    const int total = LPB * qpl;
    int l = tid / qpl, qx = tid - (tid / qpl) * qpl;
    const int dl = NT / qpl, dq = NT - (NT / qpl) * qpl;
    auto batch = [&](const int it, const int e0) {
      int li[UQ], xq[UQ];
#pragma unroll
      for (int b = 0; b < UQ; b++) {
        if constexpr (FQ) {
          const int e = e0 + b * NT + tid;
          li[b] = e / qpl;
          xq[b] = e % qpl;
        } else {
          li[b] = l;
          xq[b] = qx;
          qx += dq;
          l += dl;
          if (qx >= qpl) { qx -= qpl; l++; }
        }
      }
      double T[UQ][4], Tym[UQ][4], Typ[UQ][4], Tzm[UQ][4], Tzp[UQ][4], TL[UQ], TR[UQ];
      typename V::Raw rx[UQ][4], rxn[UQ], ry[UQ][4], ryp[UQ][4], rz[UQ][4], rzp[UQ][4];
#pragma unroll
      for (int b = 0; b < UQ; b++) {
        const int lb = FQ || li[b] < LPB ? li[b] : LPB - 1;
        const int x0 = xq[b] << 2;
        const int b0 = lbase[lb][0];
        const int xl = x0 > 0 ? x0 - 1 : 0, xr = x0 + 4 < nx ? x0 + 4 : nx - 1;
        if constexpr (PZ4) {
          if (it == 1) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
              T[b][q] = cT2[b][q];
              Tzm[b][q] = cT1[b][q];
              rz[b][q] = cRz2[b][q];
            }
          }
        }
        if constexpr (BUF) {
          // 32-bit byte offsets from buffer descriptors: no 64-bit address math
          const unsigned o[5] = {(unsigned)(b0 + x0), (unsigned)(lbase[lb][1] + x0),
                                 (unsigned)(lbase[lb][2] + x0), (unsigned)(lbase[lb][3] + x0),
                                 (unsigned)(lbase[lb][4] + x0)};
          double (*dst[5])[4] = {&T[b], &Tym[b], &Typ[b], &Tzm[b], &Tzp[b]};
#pragma unroll
          for (int k = 0; k < 5; k++) {
            if (PZ && ((b == 0 && k == 2) || (b == 1 && k == 1))) continue;  // set below
            if (PZ4 && it == 1 && (k == 0 || k == 3)) continue;              // carried
            const double2 v0 = buf_load_d2(bT, o[k] * 8u);
            const double2 v1 = buf_load_d2(bT, o[k] * 8u + 16u);
            (*dst[k])[0] = v0.x; (*dst[k])[1] = v0.y; (*dst[k])[2] = v1.x; (*dst[k])[3] = v1.y;
          }
          uint32_t wx;
          if constexpr (COMB && PZ4) wx = it == 1 ? cWz[b] : buf_load_u32(bC[2], o[0]);
          else wx = buf_load_u32(bC[2], o[0]);
          unpack4(wx, rx[b]);
          if constexpr (RY) rxw[it][b] = wx;
          cbuf[li[b] * (SEG * G / 4) + xq[b]] = wx;
          if constexpr (FQ && (SEG * G) % (4 * kWave) == 0 && NT % kWave == 0) {
            // a wave holds 64 consecutive quads of one line, quad x0/4 in lane
            // x0/4 % 64: the cells either side of the quad are in the
            // neighbouring lanes; at a wave's edge inside the line the lane
            // loads them, at the line's ends they are the cell itself, as
            // the clamped loads give
            const int lane = xq[b] & (kWave - 1);
            const double l3 = __shfl_up(T[b][3], 1, kWave);
            const double r0 = __shfl_down(T[b][0], 1, kWave);
            const uint32_t rw = __shfl_down(wx, 1, kWave);
            TL[b] = lane > 0 ? l3 : T[b][0];
            TR[b] = lane < kWave - 1 ? r0 : T[b][3];
            rxn[b] = lane < kWave - 1 ? (rw & 0xffu) : rx[b][3];
            if constexpr (SEG * G > 4 * kWave) {
              if (lane == 0 && x0 > 0) TL[b] = buf_load_d(bT, (unsigned)(b0 + x0 - 1) * 8u);
              if (lane == kWave - 1 && x0 + 4 < nx) {
                TR[b] = buf_load_d(bT, (unsigned)(b0 + x0 + 4) * 8u);
                rxn[b] = buf_load_u8(bC[2], (unsigned)(b0 + x0 + 4));
              }
            }
          } else {
            TL[b] = buf_load_d(bT, (unsigned)(b0 + xl) * 8u);
            TR[b] = buf_load_d(bT, (unsigned)(b0 + xr) * 8u);
            rxn[b] = buf_load_u8(bC[2], (unsigned)(b0 + xr));
          }
          if constexpr (COMB) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
              ry[b][q] = rx[b][q];
              rz[b][q] = rx[b][q];
            }
            if (!(PZ && b == 0)) unpack4(buf_load_u32(bC[1], o[2]), ryp[b]);
            const uint32_t wzp = buf_load_u32(bC[0], o[4]);
            unpack4(wzp, rzp[b]);
            if constexpr (PZ4) cWz[b] = wzp;
          } else {
            unpack4(buf_load_u32(bC[1], o[0]), ry[b]);
            if (!(PZ && b == 0)) unpack4(buf_load_u32(bC[1], o[2]), ryp[b]);
            if (!(PZ4 && it == 1)) unpack4(buf_load_u32(bC[0], o[0]), rz[b]);
            unpack4(buf_load_u32(bC[0], o[4]), rzp[b]);
          }
        } else {
        const double2 *p0 = reinterpret_cast<const double2 *>(Tn + b0 + x0);
        const double2 *p1 = reinterpret_cast<const double2 *>(Tn + lbase[lb][1] + x0);
        const double2 *p2 = reinterpret_cast<const double2 *>(Tn + lbase[lb][2] + x0);
        const double2 *p3 = reinterpret_cast<const double2 *>(Tn + lbase[lb][3] + x0);
        const double2 *p4 = reinterpret_cast<const double2 *>(Tn + lbase[lb][4] + x0);
        double2 v;
        v = p0[0]; T[b][0] = v.x; T[b][1] = v.y; v = p0[1]; T[b][2] = v.x; T[b][3] = v.y;
        v = p1[0]; Tym[b][0] = v.x; Tym[b][1] = v.y; v = p1[1]; Tym[b][2] = v.x; Tym[b][3] = v.y;
        v = p2[0]; Typ[b][0] = v.x; Typ[b][1] = v.y; v = p2[1]; Typ[b][2] = v.x; Typ[b][3] = v.y;
        v = p3[0]; Tzm[b][0] = v.x; Tzm[b][1] = v.y; v = p3[1]; Tzm[b][2] = v.x; Tzm[b][3] = v.y;
        v = p4[0]; Tzp[b][0] = v.x; Tzp[b][1] = v.y; v = p4[1]; Tzp[b][2] = v.x; Tzp[b][3] = v.y;
        TL[b] = Tn[b0 + xl];
        TR[b] = Tn[b0 + xr];
        view.fetch4(2, b0 + x0, rx[b]);
        rxn[b] = view.fetch(2, b0 + xr);
        view.fetch4(1, b0 + x0, ry[b]);
        view.fetch4(0, b0 + x0, rz[b]);
        if constexpr (V::kAsym) {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            ryp[b][u] = ry[b][u];
            rzp[b][u] = rz[b][u];
          }
        } else {
          view.fetch4(1, lbase[lb][2] + x0, ryp[b]);
          view.fetch4(0, lbase[lb][4] + x0, rzp[b]);
        }
        }
      }
      if constexpr (PZ) {  // quad 1 is the row after quad 0
#pragma unroll
        for (int q = 0; q < 4; q++) {
          Typ[0][q] = T[1][q];
          Tym[1][q] = T[0][q];
          ryp[0][q] = ry[1][q];
        }
      }
      if constexpr (PZ4) {  // the second batch is the plane above
        if (it == 0) {
#pragma unroll
          for (int b = 0; b < 2; b++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
              cT1[b][q] = T[b][q];
              cT2[b][q] = Tzp[b][q];
              cRz2[b][q] = rzp[b][q];
            }
        }
      }
#pragma unroll
      for (int b = 0; b < UQ; b++) {
        const int lb = FQ || li[b] < LPB ? li[b] : LPB - 1;
        const int x0 = xq[b] << 2;
        Entry ex[5];
#pragma unroll
        for (int q = 0; q < 4; q++) ex[q] = view.resolve(tabx, rx[b][q]);
        ex[4] = view.resolve(tabx, rxn[b]);
        double out[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int x = x0 + q;
          const double T0 = T[b][q];
          const double tl = q > 0 ? T[b][q - 1] : TL[b];
          const double tr = q < 3 ? T[b][q + 1] : TR[b];
          const double gxm = ex[q].g, gxp = V::kAsym ? ex[q].gp : ex[q + 1].g;
          const double gym = view.resolve(taby, ry[b][q]).g;
          const double gyp = plus_g(view, taby, ry[b][q], ryp[b][q]);
          const double gzm = view.resolve(tabz, rz[b][q]).g;
          const double gzp = plus_g(view, tabz, rz[b][q], rzp[b][q]);
          double expl;
          if constexpr (RY) {
            const double ly = gym * (Tym[b][q] - T0) + gyp * (Typ[b][q] - T0);
            expl = 0.5 * (gxm * (tl - T0) + gxp * (tr - T0)) + ly + gzm * (Tzm[b][q] - T0) +
                   gzp * (Tzp[b][q] - T0);
            hly[it][b][q] = 0.5 * ex[q].w * ly;
          } else {
            expl = 0.5 * (gxm * (tl - T0) + gxp * (tr - T0)) + gym * (Tym[b][q] - T0) +
                   gyp * (Typ[b][q] - T0) + gzm * (Tzm[b][q] - T0) + gzp * (Tzp[b][q] - T0);
          }
          if (SRC) {
            if (vmap) expl += vtab[vmap[lbase[lb][0] + x]];
            if (sfield) expl += sfield[lbase[lb][0] + x];
          }
          out[q] = ex[q].M * T0 + ex[q].w * expl;
        }
        if (FQ || li[b] < LPB) {
          const int o = li[b] * RS + x0 + (x0 >> 4);
#pragma unroll
          for (int q = 0; q < 4; q++) rbuf[o + q] = out[q];
          if constexpr (PARK) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
              pM[o + q] = rx[b][q].M;
              pG[o + q] = rx[b][q].g;
              pW[o + q] = (uint8_t)rx[b][q].w;
            }
          }
        }
      }
    };
    if constexpr (RY) {
#pragma unroll
      for (int it = 0; it < NIT; it++) batch(it, it * UQ * NT);
    } else {
#pragma unroll 1
      for (int e0 = 0; e0 < total; e0 += UQ * NT) batch(0, e0);
    }
  } else {
    // general nx: one cell per thread per item, U items per batch
    constexpr int U = 4;
    const int total = LPB * nx;
    int l = tid / nx, i = tid - (tid / nx) * nx;
    const int dl = NT / nx, di = NT - (NT / nx) * nx;
    for (int e0 = 0; e0 < total; e0 += U * NT) {
      int li[U], ii[U];
#pragma unroll
      for (int b = 0; b < U; b++) {
        li[b] = l;
        ii[b] = i;
        i += di;
        l += dl;
        if (i >= nx) { i -= nx; l++; }
      }
      double Tc[U], Txm[U], Txp[U], Tym[U], Typ[U], Tzm[U], Tzp[U];
      typename V::Raw rx[U], rxp[U], ry[U], ryp[U], rz[U], rzp[U];
#pragma unroll
      for (int b = 0; b < U; b++) {
        const int lb = li[b] < LPB ? li[b] : LPB - 1;
        const int x = ii[b];
        const int xm = x > 0 ? x - 1 : x, xp = x < nx - 1 ? x + 1 : x;
        const int b0 = lbase[lb][0];
        Tc[b] = Tn[b0 + x];
        Txm[b] = Tn[b0 + xm];
        Txp[b] = Tn[b0 + xp];
        Tym[b] = Tn[lbase[lb][1] + x];
        Typ[b] = Tn[lbase[lb][2] + x];
        Tzm[b] = Tn[lbase[lb][3] + x];
        Tzp[b] = Tn[lbase[lb][4] + x];
        rx[b] = view.fetch(2, b0 + x);
        rxp[b] = V::kAsym ? rx[b] : view.fetch(2, b0 + xp);
        ry[b] = view.fetch(1, b0 + x);
        ryp[b] = V::kAsym ? ry[b] : view.fetch(1, lbase[lb][2] + x);
        rz[b] = view.fetch(0, b0 + x);
        rzp[b] = V::kAsym ? rz[b] : view.fetch(0, lbase[lb][4] + x);
      }
#pragma unroll
      for (int b = 0; b < U; b++) {
        const int lb = li[b] < LPB ? li[b] : LPB - 1;
        const int x = ii[b];
        const Entry e = view.resolve(tabx, rx[b]);
        const double gxm = e.g;
        const double gxp = plus_g(view, tabx, rx[b], rxp[b]);
        const double gym = view.resolve(taby, ry[b]).g;
        const double gyp = plus_g(view, taby, ry[b], ryp[b]);
        const double gzm = view.resolve(tabz, rz[b]).g;
        const double gzp = plus_g(view, tabz, rz[b], rzp[b]);
        const double T0 = Tc[b];
        double expl = 0.5 * (gxm * (Txm[b] - T0) + gxp * (Txp[b] - T0)) + gym * (Tym[b] - T0) +
                      gyp * (Typ[b] - T0) + gzm * (Tzm[b] - T0) + gzp * (Tzp[b] - T0);
        if (SRC) {
          if (vmap) expl += vtab[vmap[lbase[lb][0] + x]];
          if (sfield) expl += sfield[lbase[lb][0] + x];
        }
        if (li[b] < LPB) {
          const int o = li[b] * RS + x + (x >> 4);
          rbuf[o] = e.M * T0 + e.w * expl;
          if constexpr (PARK) {
            pM[o] = rx[b].M;
            pG[o] = rx[b].g;
            pW[o] = (uint8_t)rx[b].w;
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 2: G lanes x SEG cells per line
  const int lq = tid / G, lig = tid % G;
  const bool lvalid = lok[lq];
  const int base = lbase[lq][0];
  const int i0 = lig * SEG;
  typename V::Raw rxs[SEG];
  if constexpr (PARK) {
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int x = min(i0 + q, nx - 1);
      const int o = lq * RS + x + (x >> 4);
      rxs[q] = DenseRaw{pM[o], pG[o], pW[o]};
    }
  } else if constexpr (BUF) {
    static_assert(SEG % 4 == 0, "BUF path reads class words");
    const uint32_t *cw = cbuf + lq * (SEG * G / 4) + i0 / 4;
#pragma unroll
    for (int q = 0; q < SEG; q += 4) {
      typename V::Raw t4[4];
      unpack4(cw[q / 4], t4);
#pragma unroll
      for (int u = 0; u < 4; u++) rxs[q + u] = t4[u];
    }
  } else if constexpr (SEG % 4 == 0) {
    if ((nx & 3) == 0 && i0 + SEG <= nx) {
#pragma unroll
      for (int q = 0; q < SEG; q += 4) {
        typename V::Raw t4[4];
        view.fetch4(2, base + i0 + q, t4);
#pragma unroll
        for (int u = 0; u < 4; u++) rxs[q + u] = t4[u];
      }
    } else {
#pragma unroll
      for (int q = 0; q < SEG; q++) rxs[q] = view.fetch(2, base + min(i0 + q, nx - 1));
    }
  } else {
#pragma unroll
    for (int q = 0; q < SEG; q++) rxs[q] = view.fetch(2, base + min(i0 + q, nx - 1));
  }
  if constexpr (SEG >= 2) {
    // right-hand sides stay in the lane's LDS slots (contiguous: SEG divides
    // 16), which also receive the solution; coefficients are resolved as the
    // rows stream by
    SegLds<SEG> seg;
    seg.rp = rbuf + lq * RS + i0 + (i0 >> 4);
    Entry ec = view.resolve(tabx, rxs[0]);
    const double g_after = LaneX<G>::down(ec.g, 1, lig);  // next lane's first face
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int x = i0 + q;
      Entry en;
      if (q + 1 < SEG) en = view.resolve(tabx, rxs[q + 1]);
      else en.g = g_after;
      if constexpr (FULL) {  // every lane's cells lie inside its (existing) line
        const double gxm = (q == 0 && lig == 0) ? 0.0 : ec.g;
        const double gxp = (q == SEG - 1 && lig == G - 1) ? 0.0 : V::kAsym ? ec.gp : en.g;
        const double am = -0.5 * ec.w * gxm, cp = -0.5 * ec.w * gxp;
        seg.row(q, am, ec.M - am - cp, cp, seg.rp[q]);
      } else {
        const bool valid = lvalid && x < nx;
        const double gxm = x > 0 ? ec.g : 0.0;
        const double gxp = x < nx - 1 ? (V::kAsym ? ec.gp : en.g) : 0.0;
        const double am = -0.5 * ec.w * gxm, cp = -0.5 * ec.w * gxp;
        const double rq = valid ? seg.rp[q] : 0.0;
        seg.row(q, valid ? am : 0.0, valid ? ec.M - am - cp : 1.0, valid ? cp : 0.0, rq);
      }
      ec = en;
    }
    group_solve_lds<SEG, G>(seg, lig);
  } else {
    Entry ex[SEG + 1];
#pragma unroll
    for (int q = 0; q < SEG; q++) ex[q] = view.resolve(tabx, rxs[q]);
    ex[SEG].g = LaneX<G>::down(ex[0].g, 1, lig);
    double rh[SEG];
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int x = min(i0 + q, nx - 1);
      rh[q] = rbuf[lq * RS + x + (x >> 4)];
    }
    SegOf<SEG> seg;
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int x = i0 + q;
      const bool valid = lvalid && x < nx;
      const double gxm = x > 0 ? ex[q].g : 0.0;
      const double gxp = x < nx - 1 ? (V::kAsym ? ex[q].gp : ex[q + 1].g) : 0.0;
      const double am = -0.5 * ex[q].w * gxm, cp = -0.5 * ex[q].w * gxp;
      seg.row(q, valid ? am : 0.0, valid ? ex[q].M - am - cp : 1.0, valid ? cp : 0.0,
              valid ? rh[q] : 0.0);
    }
    double xs[SEG];
    group_solve<SEG, G>(seg, xs, lig);
    // write back through LDS so that the global stores are coalesced (each
    // lane owns SEG consecutive cells; its reads of rbuf above touched only
    // those cells)
#pragma unroll
    for (int q = 0; q < SEG; q++) {
      const int x = i0 + q;
      if (x < nx) rbuf[lq * RS + x + (x >> 4)] = xs[q];
    }
  }
  __syncthreads();
  if constexpr (RY) {
    constexpr int qpl = SEG * G / 4;
#pragma unroll
    for (int it = 0; it < NIT; it++) {
#pragma unroll
      for (int b = 0; b < UQ; b++) {
        const int e = it * UQ * NT + b * NT + tid;
        const int l2 = e / qpl, x0 = (e % qpl) << 2;
        const double *src = rbuf + l2 * RS + x0 + (x0 >> 4);
        uint32_t r4[4];
        unpack4(rxw[it][b], r4);
        double o4[4];
#pragma unroll
        for (int q = 0; q < 4; q++) o4[q] = view.resolve(tabx, r4[q]).M * src[q] - hly[it][b][q];
        const unsigned o = (unsigned)(lbase[l2][0] + x0) * 8u;
        buf_store_d2(make_double2(o4[0], o4[1]), bS, o);
        buf_store_d2(make_double2(o4[2], o4[3]), bS, o + 16u);
      }
    }
  } else if ((nx & 3) == 0) {
    const int qpl = nx >> 2;
    for (int e = tid; e < LPB * qpl; e += NT) {
      const int l2 = e / qpl, x0 = (e - l2 * qpl) << 2;
      if (FULL || lok[l2]) {
        const double *src = rbuf + l2 * RS + x0 + (x0 >> 4);
        if constexpr (BUF) {
          const unsigned o = (unsigned)(lbase[l2][0] + x0) * 8u;
          buf_store_d2(make_double2(src[0], src[1]), bS, o);
          buf_store_d2(make_double2(src[2], src[3]), bS, o + 16u);
        } else {
          double2 *dst = reinterpret_cast<double2 *>(Ts + lbase[l2][0] + x0);
          dst[0] = make_double2(src[0], src[1]);
          dst[1] = make_double2(src[2], src[3]);
        }
      }
    }
  } else {
    for (int e = tid; e < LPB * nx; e += NT) {
      const int l2 = e / nx, x = e - l2 * nx;
      if (lok[l2]) Ts[lbase[l2][0] + x] = rbuf[l2 * RS + x + (x >> 4)];
    }
  }
}

template <class V, int SEG, int G, int LPB, bool SRC, bool FULL, bool RY = false, bool PZ = false>
__global__ __launch_bounds__(G *LPB) void k_sweep_x2(Grid g, V view,
                                                    const double *__restrict__ Tn,
                                                    double *__restrict__ Ts,
                                                    const uint8_t *__restrict__ vmap,
                                                    const double *__restrict__ vtab_g,
                                                    const double *__restrict__ sfield, int ty) {
  constexpr bool BUF = FULL && SEG % 4 == 0 &&
                      (std::is_same<V, ClassView>::value || std::is_same<V, CombView>::value);
  sweep_x2_body<V, SEG, G, LPB, SRC, FULL, BUF, RY, PZ>(g, view, Tn, Ts, vmap, vtab_g, sfield,
                                                        xcd_remap(blockIdx.x, gridDim.x), ty);
}

// ---------------------------------------------------------- y / z sweeps
// axis 1 (y): columns (k, i), line stride nx.  axis 0 (z): columns (j, i),
// line stride nx*ny.  Solves (M - La/2) X = M Y - (La/2) T^n along each line.
// X may alias Y or T^n: a block owns whole lines, and all of its loads
// happen before its first barrier and all of its stores after its last.
//
// FLAT (batched instances): block bx covers CW consecutive columns of all the
// instance's lines taken in memory order (axis 0: the (j, i) plane, axis 1:
// the (k, i) columns), and by is P, the segments per line; P x CW may fall
// short of the block, whose remaining threads only take part in barriers.
template <class V, int SEG, int AXIS, bool FULL = false, bool RY = false, bool FLAT = false>
__device__ __forceinline__ void sweep_strided_body(Grid g, V view, int CW, const double *Tn,
                                                   const double *Y, double *X, double *surf,
                                                   double dz, int bx, int by) {
  constexpr int axis = AXIS;
  static_assert(!(FLAT && FULL), "FLAT lines are masked");
  extern __shared__ double lds[];
  view.load_tables(lds, axis);
  const int P = FLAT ? by : blockDim.x / CW;
  double *ra = lds + ((view.lds_doubles(axis) + 1) & ~1);
  double *rc = ra + 2 * P * CW;
  double *rr = rc + 2 * P * CW;
  __syncthreads();

  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int cl = threadIdx.x % CW, p = threadIdx.x / CW;
  const bool act = !FLAT || p < P;
  const int L = axis == 1 ? ny : nz;
  const int stride = axis == 1 ? nx : nx * ny;
  bool valid;
  int base;
  if constexpr (FLAT) {
    const int ncol = nx * (axis == 1 ? nz : ny);
    const int col = bx * CW + cl;
    valid = act && col < ncol;
    const int cc = min(col, ncol - 1);
    if (axis == 1) {
      const int k = cc / nx;
      base = k * ny * nx + (cc - k * nx);
    } else {
      base = cc;
    }
  } else {
    const int i = bx * CW + cl;
    valid = i < nx;
    const int ic = valid ? i : nx - 1;
    base = axis == 1 ? by * ny * nx + ic : by * nx + ic;
  }
  const int t0 = p * SEG;

  // ---- loads: T^n over the segment plus one halo cell each side, Y, entries
  double T[SEG + 2], Yv[SEG];
  typename V::Raw raw[SEG + 1];
  if constexpr (FULL) {  // segments tile the line exactly: only the halo cells clamp
    const int b0 = base + t0 * stride;
    if constexpr (!RY) {
      T[0] = Tn[base + max(t0 - 1, 0) * stride];
#pragma unroll
      for (int q = 1; q <= SEG; q++) T[q] = Tn[b0 + (q - 1) * stride];
      T[SEG + 1] = Tn[base + min(t0 + SEG, L - 1) * stride];
    }
#pragma unroll
    for (int q = 0; q < SEG; q++) Yv[q] = Y[b0 + q * stride];
#pragma unroll
    for (int q = 0; q < SEG; q++) raw[q] = view.fetch(axis, b0 + q * stride);
    raw[SEG] = view.fetch(axis, base + min(t0 + SEG, L - 1) * stride);
  } else {
#pragma unroll
    for (int q = 0; q < SEG + 2; q++) {
      const int t = min(max(t0 + q - 1, 0), L - 1);
      T[q] = Tn[base + t * stride];
    }
#pragma unroll
    for (int q = 0; q < SEG; q++) Yv[q] = Y[base + min(t0 + q, L - 1) * stride];
#pragma unroll
    for (int q = 0; q <= SEG; q++) raw[q] = view.fetch(axis, base + min(t0 + q, L - 1) * stride);
  }

  // ---- rows, normalised to unit diagonal (independent of each other, so the
  // divisions overlap), then reduced to the segment's two end rows
  double a[SEG], c[SEG], r[SEG];
  Entry en = view.resolve(lds, raw[0]);
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    const Entry ep = view.resolve(lds, raw[q + 1]);
    // FULL: only the line's first and last cells have an outer face
    const double gm = (FULL ? (q > 0 || p > 0) : t >= 1) ? en.g : 0.0;
    const double gp =
        (FULL ? (q < SEG - 1 || t0 + SEG < L) : t + 1 <= L - 1) ? (V::kAsym ? en.gp : ep.g) : 0.0;
    const double M = en.M, w = en.w;
    // RY: Y already holds M T* - 0.5 w (Ly T^n), written by the x sweep
    const double rhs =
        RY ? Yv[q] : M * Yv[q] - 0.5 * w * (gm * (T[q] - T[q + 1]) + gp * (T[q + 2] - T[q + 1]));
    const double am = -0.5 * w * gm, cp = -0.5 * w * gp;
    const double inv = 1.0 / (M - am - cp);
    if constexpr (FULL) {
      a[q] = am * inv;
      c[q] = cp * inv;
      r[q] = rhs * inv;
    } else {
      const bool ok = valid && t < L;
      a[q] = ok ? am * inv : 0.0;
      c[q] = ok ? cp * inv : 0.0;
      r[q] = ok ? rhs * inv : 0.0;
    }
    en = ep;
  }
  // 16-cell segments are reduced by halves: two shorter division chains
  constexpr bool HALVES = SEG == 16;
  double As, Cs, Rs, Ae, Ce, Re;
  [[maybe_unused]] HalfJoin hj;
  if constexpr (HALVES)
    segment_reduce_halves<SEG>(a, c, r, As, Cs, Rs, Ae, Ce, Re, hj);
  else
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
  if constexpr (HALVES) {
    constexpr int H = SEG / 2;
    const double x1 = hj.P1 + hj.A1 * xs + hj.C1 * xe;
    const double x2 = hj.P2 + hj.A2 * xs + hj.C2 * xe;
    x[H - 1] = x1;
    x[H] = x2;
#pragma unroll
    for (int q = 1; q < H - 1; q++) x[q] = r[q] - a[q] * xs - c[q] * x1;
#pragma unroll
    for (int q = H + 1; q < SEG - 1; q++) x[q] = r[q] - a[q] * x2 - c[q] * xe;
  } else {
#pragma unroll
    for (int q = 1; q < SEG - 1; q++) x[q] = r[q] - a[q] * xs - c[q] * xe;
  }
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    const int t = t0 + q;
    if (t < L) X[base + t * stride] = x[q];
  }
  if (surf && axis == 0 && p == 0 && L >= 2) surf[base] = surface_readout(x[0], x[1], dz);
}

template <class V, int SEG, int AXIS, bool FULL, bool RY = false>
__global__ __launch_bounds__(512) void k_sweep_strided(Grid g, V view, int CW, const double *Tn,
                                                       const double *Y, double *X, double *surf,
                                                       double dz) {
  // consecutive column blocks share the 128-byte lines of the uint8 class
  // rows: keep them on one XCD (one L2), in chunks of 64 blocks (at 256^3
  // y 61.2 -> 56.0 us and z 76.5 -> 73.8 us against contiguous eighths of the
  // grid per XCD; chunks of 16 and 256 lose to 64)
  const int nb = gridDim.x * gridDim.y;
  const int r = xcd_remap_chunks<64>(blockIdx.y * gridDim.x + blockIdx.x, nb);
  sweep_strided_body<V, SEG, AXIS, FULL, RY>(g, view, CW, Tn, Y, X, surf, dz, r % gridDim.x,
                                             r / gridDim.x);
}

// ------------------------------------------------------ K0: classify
// cls_a[c] = class of cell c on axis a: lut_a[(face class*nmat + m_minus)*nmat
// + m_cell] for an interior minus face, edge_a[m_cell] on the domain edge.
__global__ void k_classify(Grid g, const uint8_t *__restrict__ mat, const uint8_t *__restrict__ bz,
                           const uint8_t *__restrict__ by, const uint8_t *__restrict__ bx,
                           const uint8_t *__restrict__ lut, const uint8_t *__restrict__ edge,
                           int nmat, int lutlen, uint8_t *cz, uint8_t *cy, uint8_t *cx) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int n = nx * ny * nz, sxy = nx * ny;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int k = c / sxy, rem = c - k * sxy, j = rem / nx, i = rem - j * nx;
    const int m = mat[c];
    cz[c] = k > 0 ? lut[0 * lutlen + (bz[c] * nmat + mat[c - sxy]) * nmat + m] : edge[0 * nmat + m];
    cy[c] = j > 0 ? lut[1 * lutlen + (by[(k * (ny + 1) + j) * nx + i] * nmat + mat[c - nx]) * nmat + m]
                  : edge[1 * nmat + m];
    cx[c] = i > 0 ? lut[2 * lutlen + (bx[(k * ny + j) * (nx + 1) + i] * nmat + mat[c - 1]) * nmat + m]
                  : edge[2 * nmat + m];
  }
}

// per-cell entries of a view as float64 arrays (for inspection and tests)
template <class V>
__global__ void k_expand(Grid g, V view, double *M, double *w, double *gz, double *gy,
                         double *gx) {
  extern __shared__ double lds[];
  view.load_tables(lds, -1);
  __syncthreads();
  const int n = g.nx * g.ny * g.nz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const Entry ez = view.resolve(view.table(lds, 0), view.fetch(0, c));
    const Entry ey = view.resolve(view.table(lds, 1), view.fetch(1, c));
    const Entry ex = view.resolve(view.table(lds, 2), view.fetch(2, c));
    if (M) M[c] = ez.M;
    if (w) w[c] = ez.w;
    if (gz) gz[c] = ez.g;
    if (gy) gy[c] = ey.g;
    if (gx) gx[c] = ex.g;
  }
}

__global__ void k_surface(const double *T, int ny, int nx, double dz, double *out) {
  const int sxy = nx * ny;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < sxy; c += gridDim.x * blockDim.x)
    out[c] = surface_readout(T[c], T[c + sxy], dz);
}

// ------------------------------------------------------------- launchers
template <class V, int CPL, int G>
static hipError_t launch_x_t(const Grid &g, const V &v, const double *Tn, double *Ts,
                             const uint8_t *vmap, const double *vtab, const double *sfield,
                             hipStream_t s) {
  const int nlines = g.nz * g.ny;
  constexpr int LPB = 256 / G;
  const int blocks = (nlines + LPB - 1) / LPB;
  const size_t shm = sizeof(double) * (((v.lds_doubles(-1) + 1) & ~1) + (vmap ? kMaxVtab : 0));
  hipLaunchKernelGGL((k_sweep_x<V, CPL, G>), dim3(blocks), dim3(256), shm, s, g, v, Tn, Ts, vmap,
                     vtab, sfield);
  return hipGetLastError();
}

// FULL two-phase launch: lines fill the lanes exactly, every block holds
// existing lines (whole tiles), and the field fits 32-bit byte offsets
static bool x2_full(const Grid &g, int lanes_cells, int lpb, int ty) {
  const int nlines = g.nz * g.ny;
  const bool whole = ty > 0 ? g.nz % (lpb / ty) == 0 && g.ny % ty == 0 : nlines % lpb == 0;
  return g.nx == lanes_cells && whole &&
         (size_t)nlines * g.nx * sizeof(double) < (size_t(1) << 32);
}

bool x_ry_ok(XConfig c, const Grid &g) {
  return c.lpb > 0 && c.cpl % 4 == 0 && x2_full(g, c.cpl * c.g, c.lpb, c.ty);
}

template <class V, int SEG, int G, int LPB>
static hipError_t launch_x2_t(const Grid &g, const V &v, int ty, const double *Tn, double *Ts,
                              const uint8_t *vmap, const double *vtab, const double *sfield,
                              bool ry, hipStream_t s) {
  const int nlines = g.nz * g.ny;
  const int tz = ty > 0 ? LPB / ty : 0;
  const int blocks = ty > 0 ? ((g.nz + tz - 1) / tz) * ((g.ny + ty - 1) / ty)
                            : (nlines + LPB - 1) / LPB;
  constexpr int RS = G * SEG + ((G * SEG) >> 4) + 1;
  const bool src = vmap || sfield;
  // BUF variants (class maps, SEG % 4 == 0) also park one class word per quad,
  // per-voxel variants each cell's x entry (M, g, w)
  const size_t shm =
      sizeof(double) * (((v.lds_doubles(-1) + 1) & ~1) + (src ? kMaxVtab : 0) + LPB * RS) +
      (std::is_same<V, ClassView>::value && SEG % 4 == 0 ? (size_t)LPB * SEG * G : 0) +
      (std::is_same<V, DenseView>::value ? (2 * sizeof(double) + 1) * (size_t)LPB * RS : 0);
  const bool full = x2_full(g, SEG * G, LPB, ty);
  if constexpr (std::is_same<V, ClassView>::value && SEG % 4 == 0) {
    // 16-cell lanes in blocks of 4, 8 or 16 lines as tiles of LPB/2 planes x
    // 2 rows: the paired line order (source-free steps on full lines always
    // take this R_y path: x_ry_ok holds exactly when the lines are full)
    constexpr bool kPZ = SEG == 16 && LPB % 4 == 0;
    if (ry) {  // the caller checked x_ry_ok and that the step has no sources
      if (!full || src) return hipErrorInvalidValue;
      if (kPZ && ty == 2)
        hipLaunchKernelGGL((k_sweep_x2<V, SEG, G, LPB, false, true, true, kPZ>), dim3(blocks),
                           dim3(G * LPB), shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
      else
        hipLaunchKernelGGL((k_sweep_x2<V, SEG, G, LPB, false, true, true>), dim3(blocks),
                           dim3(G * LPB), shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
      return hipGetLastError();
    }
  }
  if (ry) return hipErrorInvalidValue;
  if (src)
    hipLaunchKernelGGL((k_sweep_x2<V, SEG, G, LPB, true, false>), dim3(blocks), dim3(G * LPB), shm,
                       s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  else if (full)
    hipLaunchKernelGGL((k_sweep_x2<V, SEG, G, LPB, false, true>), dim3(blocks), dim3(G * LPB), shm,
                       s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  else
    hipLaunchKernelGGL((k_sweep_x2<V, SEG, G, LPB, false, false>), dim3(blocks), dim3(G * LPB),
                       shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  return hipGetLastError();
}

XConfig x_config_for(int nx, int nz) {
  // two-phase kernel, 16 lanes per line where a line fits 16 x 16 cells;
  // blocks of 4 or more lines as tiles 2 rows wide (when there are planes
  // enough to stack them)
  XConfig c{0, 0, 0, 0};
  if (nx <= 16) c = {1, 16, 8};
  else if (nx <= 32) c = {2, 16, 8};
  else if (nx <= 64) c = {4, 16, 8};
  else if (nx <= 128) c = {8, 16, 8};
  else if (nx <= 256) c = {16, 16, 16};
  else if (nx <= 512) c = {16, 32, 4};
  else if (nx <= 1024) c = {16, 64, 2};
  if (c.lpb >= 4 && 2 * nz >= c.lpb) c.ty = 2;
  return c;
}

template <class V>
static hipError_t launch_x_v(XConfig cfg, const Grid &g, const V &v, const double *Tn,
                             double *Ts, const uint8_t *vmap, const double *vtab,
                             const double *sfield, bool ry, hipStream_t s) {
  if (ry && cfg.lpb == 0) return hipErrorInvalidValue;
#define HS2_X(C, GG)                                                               \
  if (cfg.lpb == 0 && cfg.cpl == C && cfg.g == GG)                                 \
    return launch_x_t<V, C, GG>(g, v, Tn, Ts, vmap, vtab, sfield, s);
  HS2_X(1, 16) HS2_X(2, 16) HS2_X(4, 16) HS2_X(8, 16) HS2_X(16, 16)
  HS2_X(2, 32) HS2_X(4, 32) HS2_X(8, 32) HS2_X(16, 32)
  HS2_X(2, 64) HS2_X(4, 64) HS2_X(8, 64) HS2_X(16, 64)
#undef HS2_X
#define HS2_X2(C, GG, LL)                                                          \
  if (cfg.lpb == LL && cfg.cpl == C && cfg.g == GG)                                \
    return launch_x2_t<V, C, GG, LL>(g, v, cfg.ty, Tn, Ts, vmap, vtab, sfield, ry, s);
  HS2_X2(1, 16, 8) HS2_X2(2, 16, 8) HS2_X2(4, 16, 8) HS2_X2(8, 16, 8) HS2_X2(16, 16, 8)
  HS2_X2(4, 16, 16) HS2_X2(8, 16, 16) HS2_X2(16, 16, 4) HS2_X2(16, 16, 16)
  HS2_X2(8, 32, 4) HS2_X2(16, 32, 4) HS2_X2(16, 32, 8) HS2_X2(16, 32, 2) HS2_X2(16, 64, 2)
  HS2_X2(8, 64, 2) HS2_X2(16, 64, 4)
#undef HS2_X2
  return hipErrorInvalidValue;
}

// x sweep on the combined class plane: the full-line buffer kernels only
// (source-free steps; x_ry_ok(cfg) holds), with or without R_y
template <int SEG, int G, int LPB>
static hipError_t launch_x2c_t(const Grid &g, const CombView &v, int ty, const double *Tn,
                               double *Ts, bool ry, hipStream_t s) {
  const int tz = ty > 0 ? LPB / ty : 0;
  const int blocks = ty > 0 ? ((g.nz + tz - 1) / tz) * ((g.ny + ty - 1) / ty)
                            : (g.nz * g.ny + LPB - 1) / LPB;
  constexpr int RS = G * SEG + ((G * SEG) >> 4) + 1;
  const size_t shm = sizeof(double) * (((v.lds_doubles(-1) + 1) & ~1) + LPB * RS) +
                     (size_t)LPB * SEG * G;
  if (!x2_full(g, SEG * G, LPB, ty)) return hipErrorInvalidValue;
  constexpr bool kPZ = SEG == 16 && LPB % 4 == 0;
  const uint8_t *vmap = nullptr;
  const double *vtab = nullptr, *sfield = nullptr;
  if (ry && kPZ && ty == 2)
    hipLaunchKernelGGL((k_sweep_x2<CombView, SEG, G, LPB, false, true, true, kPZ>), dim3(blocks),
                       dim3(G * LPB), shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  else if (ry)
    hipLaunchKernelGGL((k_sweep_x2<CombView, SEG, G, LPB, false, true, true>), dim3(blocks),
                       dim3(G * LPB), shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  else
    hipLaunchKernelGGL((k_sweep_x2<CombView, SEG, G, LPB, false, true>), dim3(blocks),
                       dim3(G * LPB), shm, s, g, v, Tn, Ts, vmap, vtab, sfield, ty);
  return hipGetLastError();
}

hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const CombView &v, const double *Tn,
                          double *Ts, hipStream_t s, bool ry) {
#define HS2_XC(C, GG, LL)                                                          \
  if (cfg.lpb == LL && cfg.cpl == C && cfg.g == GG)                                \
    return launch_x2c_t<C, GG, LL>(g, v, cfg.ty, Tn, Ts, ry, s);
  HS2_XC(4, 16, 8) HS2_XC(8, 16, 8) HS2_XC(16, 16, 8) HS2_XC(4, 16, 16) HS2_XC(8, 16, 16)
  HS2_XC(16, 16, 4) HS2_XC(16, 16, 16) HS2_XC(8, 32, 4) HS2_XC(16, 32, 4) HS2_XC(16, 32, 8)
  HS2_XC(16, 32, 2) HS2_XC(16, 64, 2) HS2_XC(8, 64, 2) HS2_XC(16, 64, 4)
#undef HS2_XC
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const ClassView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry) {
  return launch_x_v(cfg, g, v, Tn, Ts, vmap, vtab, sfield, ry, s);
}
hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const DenseView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry) {
  return launch_x_v(cfg, g, v, Tn, Ts, vmap, vtab, sfield, ry, s);
}
hipError_t launch_sweep_x(XConfig cfg, const Grid &g, const AsymView &v, const double *Tn,
                          double *Ts, const uint8_t *vmap, const double *vtab,
                          const double *sfield, hipStream_t s, bool ry) {
  return launch_x_v(cfg, g, v, Tn, Ts, vmap, vtab, sfield, ry, s);
}

template <class V>
static size_t strided_lds_bytes_t(const V &v, int axis, int P, int CW) {
  return sizeof(double) * ((size_t)((v.lds_doubles(axis) + 1) & ~1) + 3 * 2 * (size_t)P * CW);
}
size_t strided_lds_bytes(const ClassView &v, int axis, int P, int CW) {
  return strided_lds_bytes_t<ClassView>(v, axis, P, CW);
}
size_t strided_lds_bytes(const DenseView &v, int axis, int P, int CW) {
  return strided_lds_bytes_t<DenseView>(v, axis, P, CW);
}

template <class V, int SEG>
static hipError_t launch_s_t(const Grid &g, const V &v, int axis, int CW, const double *Tn,
                             const double *Y, double *X, double *surf, double dz, bool ry,
                             hipStream_t s) {
  const int L = axis == 1 ? g.ny : g.nz;
  const int P = (L + SEG - 1) / SEG;
  const int other = axis == 1 ? g.nz : g.ny;
  dim3 grid((g.nx + CW - 1) / CW, other);
  const size_t lds = strided_lds_bytes_t<V>(v, axis, P, CW);
  // FULL: the segments tile the line and the column blocks tile nx
  const bool full = P * SEG == L && g.nx % CW == 0;
  const dim3 block(P * CW);
  if (ry) {  // y from the x sweep's R_y (class maps only)
    if constexpr (std::is_same<V, ClassView>::value) {
      if (axis != 1) return hipErrorInvalidValue;
      if (full)
        hipLaunchKernelGGL((k_sweep_strided<V, SEG, 1, true, true>), grid, block, lds, s, g, v, CW,
                           Tn, Y, X, surf, dz);
      else
        hipLaunchKernelGGL((k_sweep_strided<V, SEG, 1, false, true>), grid, block, lds, s, g, v,
                           CW, Tn, Y, X, surf, dz);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (axis == 1) {
    if (full)
      hipLaunchKernelGGL((k_sweep_strided<V, SEG, 1, true>), grid, block, lds, s, g, v, CW, Tn, Y,
                         X, surf, dz);
    else
      hipLaunchKernelGGL((k_sweep_strided<V, SEG, 1, false>), grid, block, lds, s, g, v, CW, Tn,
                         Y, X, surf, dz);
  } else {
    if (full)
      hipLaunchKernelGGL((k_sweep_strided<V, SEG, 0, true>), grid, block, lds, s, g, v, CW, Tn, Y,
                         X, surf, dz);
    else
      hipLaunchKernelGGL((k_sweep_strided<V, SEG, 0, false>), grid, block, lds, s, g, v, CW, Tn,
                         Y, X, surf, dz);
  }
  return hipGetLastError();
}

template <class V>
static hipError_t launch_s_v(SConfig cfg, const Grid &g, const V &v, int axis, const double *Tn,
                             const double *Y, double *X, double *surf, double dz, bool ry,
                             hipStream_t s) {
  switch (cfg.seg) {
    case 2: return launch_s_t<V, 2>(g, v, axis, cfg.cw, Tn, Y, X, surf, dz, ry, s);
    case 4: return launch_s_t<V, 4>(g, v, axis, cfg.cw, Tn, Y, X, surf, dz, ry, s);
    case 8: return launch_s_t<V, 8>(g, v, axis, cfg.cw, Tn, Y, X, surf, dz, ry, s);
    case 16: return launch_s_t<V, 16>(g, v, axis, cfg.cw, Tn, Y, X, surf, dz, ry, s);
    case 32: return launch_s_t<V, 32>(g, v, axis, cfg.cw, Tn, Y, X, surf, dz, ry, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const ClassView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry) {
  return launch_s_v(cfg, g, v, axis, Tn, Y, X, surf, dz, ry, s);
}
hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const DenseView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry) {
  return launch_s_v(cfg, g, v, axis, Tn, Y, X, surf, dz, ry, s);
}
hipError_t launch_sweep_strided(SConfig cfg, const Grid &g, const AsymView &v, int axis,
                                const double *Tn, const double *Y, double *X, double *surf,
                                double dz, hipStream_t s, bool ry) {
  return launch_s_v(cfg, g, v, axis, Tn, Y, X, surf, dz, ry, s);
}

// triple key of a cell: z class | y class << 8 | x class << 16
__global__ void k_comb_mark(int n, const uint8_t *__restrict__ cz, const uint8_t *__restrict__ cy,
                            const uint8_t *__restrict__ cx, uint32_t *bits) {
  // few distinct keys over many cells: a wave whose lanes share one key sets
  // it once, and a bit already seen set is not set again (a stale read only
  // costs an extra atomic), so the atomics do not pile up on a few words
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const uint32_t key = cz[c] | ((uint32_t)cy[c] << 8) | ((uint32_t)cx[c] << 16);
    const uint32_t first = __shfl(key, __builtin_ctzll(__ballot(1)), kWave);
    const bool one = __all(key == first);
    const bool lead = (threadIdx.x & (kWave - 1)) == (unsigned)__builtin_ctzll(__ballot(1));
    if ((!one || lead) && !(bits[key >> 5] & (1u << (key & 31))))
      atomicOr(bits + (key >> 5), 1u << (key & 31));
  }
}

// comb[c] = rank of the cell's key among the marked keys (rank0[w]: marked
// keys below word w); the caller checked that there are 256 or fewer
__global__ void k_comb_assign(int n, const uint8_t *__restrict__ cz,
                              const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cx,
                              const uint32_t *__restrict__ bits, const int *__restrict__ rank0,
                              uint8_t *comb) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const uint32_t key = cz[c] | ((uint32_t)cy[c] << 8) | ((uint32_t)cx[c] << 16);
    const uint32_t below = bits[key >> 5] & ((1u << (key & 31)) - 1u);
    comb[c] = (uint8_t)(rank0[key >> 5] + __popc(below));
  }
}

hipError_t launch_comb_mark(int n, const uint8_t *cz, const uint8_t *cy, const uint8_t *cx,
                            uint32_t *bits, hipStream_t s) {
  const int blocks = std::min((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_comb_mark, dim3(blocks), dim3(256), 0, s, n, cz, cy, cx, bits);
  return hipGetLastError();
}

hipError_t launch_comb_assign(int n, const uint8_t *cz, const uint8_t *cy, const uint8_t *cx,
                              const uint32_t *bits, const int *rank0, uint8_t *comb,
                              hipStream_t s) {
  const int blocks = std::min((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_comb_assign, dim3(blocks), dim3(256), 0, s, n, cz, cy, cx, bits, rank0,
                     comb);
  return hipGetLastError();
}

hipError_t launch_classify(const Grid &g, const uint8_t *mat, const uint8_t *bz, const uint8_t *by,
                           const uint8_t *bx, const uint8_t *lut, const uint8_t *edge, int nmat,
                           int lutlen, uint8_t *cz, uint8_t *cy, uint8_t *cx, hipStream_t s) {
  const int n = g.nz * g.ny * g.nx;
  int blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(256), 0, s, g, mat, bz, by, bx, lut, edge,
                     nmat, lutlen, cz, cy, cx);
  return hipGetLastError();
}

template <class V>
static hipError_t launch_expand_v(const Grid &g, const V &v, double *M, double *w, double *gz,
                                  double *gy, double *gx, hipStream_t s) {
  const int n = g.nz * g.ny * g.nx;
  int blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_expand<V>), dim3(blocks), dim3(256), sizeof(double) * v.lds_doubles(-1), s,
                     g, v, M, w, gz, gy, gx);
  return hipGetLastError();
}
hipError_t launch_expand(const Grid &g, const ClassView &v, double *M, double *w, double *gz,
                         double *gy, double *gx, hipStream_t s) {
  return launch_expand_v(g, v, M, w, gz, gy, gx, s);
}
hipError_t launch_expand(const Grid &g, const DenseView &v, double *M, double *w, double *gz,
                         double *gy, double *gx, hipStream_t s) {
  return launch_expand_v(g, v, M, w, gz, gy, gx, s);
}

hipError_t launch_surface(const double *T, int ny, int nx, double dz, double *out, hipStream_t s) {
  const int sxy = nx * ny;
  int blocks = (sxy + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(256), 0, s, T, ny, nx, dz, out);
  return hipGetLastError();
}


// ------------------------------------------------------ batched sweeps
// One launch covers every instance of a batch (per-voxel coefficients,
// axis lengths <= kBatchMaxLen); see hs2_batch.h for the block mapping.
// Lines of T^t of one x-sweep block (LPB whole x-lines, contiguous) copied
// into the instance's checkpoint slot when step t starts a checkpoint
// segment: (nsteps - t) % K == 0, or t == 0 (the segments are aligned to
// the instance's end).
__device__ __forceinline__ void batch_ckpt_lines(const BatchMap &m, const BatchInst &I, int t,
                                                 int local, int lpb, const double *T,
                                                 double *ck) {
  const int r = I.nsteps - t;
  if (r % m.K != 0 && t != 0) return;
  const long long n = (long long)I.nz * I.ny * I.nx;
  const int slot = (r + m.K - 1) / m.K - 1;
  const long long c0 = (long long)local * lpb * I.nx;
  const long long c1 = min(n, c0 + (long long)lpb * I.nx);
  double *dst = ck + I.ckoff + (long long)slot * n;
  for (long long c = c0 + threadIdx.x; c < c1; c += blockDim.x) dst[c] = T[I.off + c];
}

template <int G, int LPB>
__global__ __launch_bounds__(G *LPB) void kb_sweep_x2(BatchMap m, DenseView v, const double *Tn,
                                                     double *Ts, const double *src, double *ck) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;
  if (ck) batch_ckpt_lines(m, I, t, local, LPB, Tn, ck);
  const double *sf = (src && t == I.src_step) ? src + I.off : nullptr;
  const Grid g{I.nz, I.ny, I.nx};
  const DenseView vi = offset_view(v, I.off);
  // lanes hold 2, 4 or 8 cells: 16 lanes cover the instance's x lines
  switch (I.xseg) {
    case 2:
      sweep_x2_body<DenseView, 2, G, LPB, true>(g, vi, Tn + I.off, Ts + I.off, nullptr, nullptr,
                                                sf, local, 0);
      break;
    case 4:
      sweep_x2_body<DenseView, 4, G, LPB, true>(g, vi, Tn + I.off, Ts + I.off, nullptr, nullptr,
                                                sf, local, 0);
      break;
    default:
      sweep_x2_body<DenseView, 8, G, LPB, true>(g, vi, Tn + I.off, Ts + I.off, nullptr, nullptr,
                                                sf, local, 0);
      break;
  }
}

template <int SEG, int AXIS>
__global__ __launch_bounds__(kBatchThreads) void kb_sweep_strided(
    BatchMap m, DenseView v, const double *Tn, const double *Y, double *X, double *hist) {
  BatchInst I;
  int local, t;
  if (!batch_locate(m, I, local, t)) return;
  double *surf = (AXIS == 0 && hist) ? hist + I.hoff + (long long)t * I.ny * I.nx : nullptr;
  sweep_strided_body<DenseView, SEG, AXIS, false, false, true>(
      Grid{I.nz, I.ny, I.nx}, offset_view(v, I.off), I.cw[AXIS], Tn + I.off, Y + I.off,
      X + I.off, surf, I.dz, local, batch_line_segs(AXIS == 1 ? I.ny : I.nz));
}

hipError_t launch_batch_x2(const BatchMap &m, int nblocks, const DenseView &v, const double *Tn,
                           double *Ts, const double *src, double *ck, hipStream_t s) {
  constexpr int G = 16, LPB = kBatchXLines, SEGMAX = kBatchMaxLen / 16;
  constexpr int RS = G * SEGMAX + ((G * SEGMAX) >> 4) + 1;
  // vtab, right-hand sides, parked x entries (M, g as doubles, w as bytes)
  const size_t shm = sizeof(double) * (kMaxVtab + 3 * LPB * RS) + (size_t)LPB * RS;
  hipLaunchKernelGGL((kb_sweep_x2<G, LPB>), dim3(nblocks), dim3(G * LPB), shm, s, m, v, Tn, Ts,
                     src, ck);
  return hipGetLastError();
}

hipError_t launch_batch_strided(const BatchMap &m, int nblocks, int axis, const DenseView &v,
                                const double *Tn, const double *Y, double *X, double *hist,
                                hipStream_t s) {
  const size_t shm = sizeof(double) * 6 * (size_t)kBatchThreads;
  if (axis == 1)
    hipLaunchKernelGGL((kb_sweep_strided<kBatchSeg, 1>), dim3(nblocks), dim3(kBatchThreads), shm,
                       s, m, v, Tn, Y, X, hist);
  else
    hipLaunchKernelGGL((kb_sweep_strided<kBatchSeg, 0>), dim3(nblocks), dim3(kBatchThreads), shm,
                       s, m, v, Tn, Y, X, hist);
  return hipGetLastError();
}

}  // namespace hs2
