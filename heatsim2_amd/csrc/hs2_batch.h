// Batched execution of many independent ADI problems of different sizes and
// step counts: one launch per sweep covers every instance.  Each launch kind
// has a block -> instance table; a block finds its instance, its local block
// index and its instance's current step, and works on that instance's slice
// of the concatenated fields (per-voxel coefficients, DenseView).
//
// Time is driven by counters in device memory (ctr[0]: forward step,
// ctr[1]: backward segment), so a captured graph of a chunk of steps can be
// replayed without changing any kernel argument.  Checkpoints are aligned
// to each instance's end: with K steps per segment, backward segment s of
// instance i covers steps [max(0, e-K), e), e = nsteps_i - s*K, so every
// instance is at the same position inside its segment and the recomputation
// runs in lockstep.  Segment slot q (0..K) holds T^{e-K+q}; T* and T** of
// step e-K+q are kept in slot q of their own stores.
#pragma once
#include <hip/hip_runtime.h>

#include "hs2_device.h"

namespace hs2 {

constexpr int kBatchMaxLen = 128;  // axis length limit of the batched kernels
constexpr int kBatchThreads = 256; // threads per block of the strided / line kernels
constexpr int kBatchXThreads = 128; // threads per block of the adjoint x-line kernel
constexpr int kBatchSeg = 8;       // cells per thread segment of the line kernels
constexpr int kBatchXLines = 8;    // x-lines per block of the batched x sweep (16 lanes each)
constexpr int kBatchAlign = 32;    // instance offsets are multiples of this many cells

// Segments of kBatchSeg cells per line of L cells in the y and z line
// kernels (the fewest that cover L), and the lines a 256-thread block holds
// side by side, cw = floor(256 / P): consecutive columns in memory order.
__host__ __device__ inline int batch_line_segs(int L) { return (L + kBatchSeg - 1) / kBatchSeg; }
__host__ __device__ inline int batch_flat_cw(int L) { return kBatchThreads / batch_line_segs(L); }
// Lines per block of the adjoint x-line kernel (threads / a power-of-two
// count of segments covering L).
__host__ __device__ inline int batch_line_cw(int L, int threads = kBatchThreads) {
  int p = 1;
  while (p * kBatchSeg < L) p <<= 1;
  return threads / p;
}
// cells per lane (16 lanes per line) of the batched x sweep's solve
__host__ __device__ inline int batch_x_seg(int nx) { return nx <= 32 ? 2 : nx <= 64 ? 4 : 8; }

struct BatchInst {
  int nz, ny, nx, nsteps;
  long long off;    // first cell in the concatenated fields
  long long hoff;   // first value of the surface history (nsteps x ny x nx)
  long long ckoff;  // first value of the checkpoints (ceil(nsteps/K) fields)
  int src_step;     // step whose source field applies (-1: none)
  int xseg;         // batch_x_seg(nx)
  int cw[3];        // lines per block: batch_flat_cw of the z and y line
                    // lengths, batch_line_cw of x's for the adjoint x-line
                    // kernel's kBatchXThreads
  int pad;
  double dz;
};

enum BatchRule { RULE_ALL = 0, RULE_FWD = 1, RULE_RECOMP = 2, RULE_BWD = 3, RULE_SEG = 4 };

struct BatchMap {
  const BatchInst *inst;
  const int *blk_inst;  // instance of each block of this launch kind
  const int *blk0;      // first block of each instance
  const int *ctr;       // device counters
  int rule, j, K;
};

// Logical block of hardware block b among nb.  Hardware hands block b to XCD
// b % 8, each with its own L2; consecutive logical blocks touch neighbouring
// rows (and share cache lines of rows that do not start on a line), so they
// go to one XCD in chunks of kBatchChunk (32: measured best among 1 to
// 1024), and the chunks go round the XCDs so
// that every XCD keeps a share of every instance (instances finish at
// different steps).  Blocks past the last whole round of chunks keep their
// index.
constexpr int kBatchChunk = 32;
__device__ __forceinline__ int batch_block(int b, int nb) {
  return xcd_remap_chunks<kBatchChunk>(b, nb);
}

// Instance, local block and step of this block; false if the instance has
// nothing to do under the launch's rule (the whole block returns).
__device__ __forceinline__ bool batch_locate(const BatchMap &m, BatchInst &I, int &local, int &t) {
  const int gb = batch_block(blockIdx.x, gridDim.x);
  const int i = m.blk_inst[gb];
  local = gb - m.blk0[i];
  I = m.inst[i];
  if (m.rule == RULE_ALL) {
    t = 0;
    return true;
  }
  if (m.rule == RULE_FWD) {
    t = m.ctr[0];
    return t < I.nsteps;
  }
  const int e = I.nsteps - m.ctr[1] * m.K;
  if (e <= 0) return false;
  const int s = e - m.K > 0 ? e - m.K : 0;
  if (m.rule == RULE_SEG) {
    t = s;
    return true;
  }
  if (m.rule == RULE_RECOMP) {  // step t: T^t (slot j) -> T^{t+1} (slot j+1)
    t = e - m.K + m.j;
    return t >= s;
  }
  t = e - 1 - m.j;  // RULE_BWD
  return t >= s;
}

__device__ __forceinline__ DenseView offset_view(DenseView v, long long off) {
  v.M += off;
  v.w += off;
  v.g[0] += off;
  v.g[1] += off;
  v.g[2] += off;
  return v;
}

}  // namespace hs2

namespace hs2 {
struct AdjFields;
// launchers (hs2_kernels.hip, hs2_adjoint.hip); nblocks = this kind's block total
// the x sweep; with ck != nullptr (forward) a block first copies its lines of
// T^t into the instance's checkpoint slot on a checkpoint step
hipError_t launch_batch_x2(const BatchMap &m, int nblocks, const DenseView &v, const double *Tn,
                           double *Ts, const double *src, double *ck, hipStream_t s);
hipError_t launch_batch_strided(const BatchMap &m, int nblocks, int axis, const DenseView &v,
                                const double *Tn, const double *Y, double *X, double *hist,
                                hipStream_t s);
hipError_t launch_batch_adj_line(const BatchMap &m, int nblocks, int axis, const DenseView &v,
                                 const double *R, int scale_m, double *U, hipStream_t s);
hipError_t launch_batch_adj_stencil(const BatchMap &m, int nblocks, const DenseView &v,
                                    const AdjFields &f, hipStream_t s);
hipError_t launch_batch_ckpt_load(const BatchMap &m, int nblocks, const double *ck, double *seg,
                                  long long N, hipStream_t s);
hipError_t launch_batch_lam_init(const BatchMap &m, int nblocks, const double *ghist, double *lam,
                                 hipStream_t s);
hipError_t launch_batch_counter(int *ctr, int which, int value, int add, hipStream_t s);
}  // namespace hs2
