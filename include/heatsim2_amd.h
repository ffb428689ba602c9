/* heatsim2_amd: C ABI of the MI355X (gfx950) 3-D ADI Crank-Nicolson step.
 *
 * This is the boundary the Python package heatsim2_amd binds (ctypes), and
 * that any other host would bind.  Conventions:
 *   - every function returns an int status (HS2_OK = 0) and never exits or
 *     aborts; hs2_last_error() gives a message for the calling thread;
 *   - every large buffer is caller-owned device memory passed as a plain
 *     pointer (fields are (nz,ny,nx) C-order float64, 16-byte aligned; maps
 *     are uint8 in the upstream layout); the plan owns its tables, the
 *     per-cell coefficient classes and one n-double workspace;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *   - a plan is used from one stream at a time; distinct plans are
 *     independent.
 *
 * Upstream interface each entry point replaces is cited as file:line of the
 * heatsim2 source tree.
 */
#ifndef HEATSIM2_AMD_H
#define HEATSIM2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HS2_OK = 0,
  HS2_E_ARG = 1,          /* bad argument (null pointer, size, index)           */
  HS2_E_UNSUPPORTED = 2,  /* valid but outside what the kernels handle         */
  HS2_E_HIP = 3,          /* a HIP runtime call failed                         */
  HS2_E_STATE = 4,        /* plan not ready (coefficients not set)             */
  HS2_E_ALLOC = 5         /* device or host allocation failed                  */
};

typedef struct hs2_plan hs2_plan;

typedef struct {
  int32_t nz, ny, nx;     /* cells along z (slowest), y, x (contiguous)       */
  double dz, dy, dx;      /* spacing, m                                        */
  double dt;              /* time step, s                                      */
} hs2_grid_desc;

/* Create a plan for one grid.
 * Replaces adi_setup + pyadi_step.__init__ / create_adi_step
 * (alternatingdirection_c_pyx.pyx:418-445, :137-167;
 *  alternatingdirection_c.h:51).  Every axis length must be <= 1024 and the
 * face-map sizes below 2^31 (HS2_E_UNSUPPORTED otherwise): the line kernels
 * hold a line in registers and LDS, and index with 32 bits. */
int hs2_plan_create(const hs2_grid_desc *grid, hs2_plan **out);

/* Release a plan.  Replaces delete_adi_step (alternatingdirection_c.h:50,
 * alternatingdirection_c.c:14-54) and pyadi_step.__del__
 * (alternatingdirection_c_pyx.pyx:169-175). */
void hs2_plan_destroy(hs2_plan *plan);

/* Compact coefficients: the upstream uint8 maps stay on the device and the
 * sweeps derive every coefficient on the fly from three small tables.
 *   d_mat (nz,ny,nx), d_bz (nz+1,ny,nx), d_by (nz,ny+1,nx), d_bx (nz,ny,nx+1):
 *       device uint8 maps, caller-owned, must outlive the plan's use;
 *   h_gtab[((b*3 + axis)*nmat + m_minus)*nmat + m_plus], axis 0=z 1=y 2=x:
 *       face conductance per volume (W m^-3 K^-1) of a class-b face between
 *       a minus-side cell of material m_minus and a plus-side cell of
 *       material m_plus, as seen by the COMPUTE cell(s) beside it (the
 *       neighbour-k fallback already applied);
 *   h_mtab[nmat]: rho*c/dt for COMPUTE materials, 1 for FIXED;
 *   h_compute[nmat]: 1 COMPUTE, 0 FIXED.
 * Outer faces carry no flux (validated by the caller, as the upstream
 * requires).  Replaces the per-cell add_equation / finalize assembly
 * (crank_nicolson.pyx:272-496; alternatingdirection_c.h:53;
 *  alternatingdirection_c_pyx.pyx:177-282). */
int hs2_plan_set_compact(hs2_plan *plan, const uint8_t *d_mat, const uint8_t *d_bz,
                         const uint8_t *d_by, const uint8_t *d_bx, int nmat, int nb,
                         const double *h_gtab, const double *h_mtab, const uint8_t *h_compute);

/* Per-voxel coefficients instead of the maps: caller-owned device arrays of
 * n entries, cell-indexed in C order -- d_M: rho*c/dt (1 for a FIXED cell);
 * d_w: 1 COMPUTE / 0 FIXED (uint8); d_gz, d_gy, d_gx: conductance per
 * volume of each cell's minus face on that axis (the face towards k-1, j-1,
 * i-1; 0 on the outer faces).  The plus face of a cell is the minus face of
 * its neighbour.  Same numbers as hs2_plan_set_compact derives, for
 * per-voxel material fields that no uint8 map can express; the upstream has
 * no per-voxel mode (its maps cap materials at 256, crank_nicolson.pyx:135). */
int hs2_plan_set_dense(hs2_plan *plan, const double *d_M, const uint8_t *d_w, const double *d_gz,
                       const double *d_gy, const double *d_gx);

/* Per-voxel coefficients whose faces are seen differently from the two
 * sides: as hs2_plan_set_dense, plus each cell's own plus-face conductances
 * d_gzp, d_gyp, d_gxp (the face towards k+1, j+1, i+1), which need not equal
 * the next cell's minus-face values.  This is the upstream's top-surface
 * curvature mode, where each cell weighs its faces by its own areas and
 * volume (crank_nicolson.pyx:388-457, replacing the per-cell add_equation
 * assembly of alternatingdirection_c.c:102-199 for that mode).  Forward
 * steps only: the gradient entry points reject such a plan. */
int hs2_plan_set_dense_asym(hs2_plan *plan, const double *d_M, const uint8_t *d_w,
                            const double *d_gzm, const double *d_gym, const double *d_gxm,
                            const double *d_gzp, const double *d_gyp, const double *d_gxp);

/* Write the plan's per-cell coefficients as float64 arrays (n each; any
 * pointer may be NULL): M, w, and the minus-face conductances per axis --
 * the numbers the upstream keeps in Amat/Bmat
 * (alternatingdirection_c_pyx.pyx:230,251), for inspection. */
int hs2_plan_expand(hs2_plan *plan, double *d_M, double *d_w, double *d_gz, double *d_gy,
                    double *d_gx, void *stream);

/* Kernel launch shapes (used by the autotuner).  x sweep: x_lanes lanes of
 * x_cpl consecutive cells per x-line (x_cpl*x_lanes >= nx; x_lanes in
 * {16,32,64}, x_cpl in {1,2,4,8,16}); x_lines > 0 selects the two-phase
 * kernel with x_lines lines per block (right-hand sides staged in LDS),
 * x_lines = 0 the single-phase one.  x_lines_y (two-phase only): 0 puts a
 * block's lines one after the other in (z, y) order; a proper divisor of
 * x_lines tiles them as x_lines/x_lines_y planes x x_lines_y rows, which
 * share more neighbour rows.  y and z sweeps: segments of *_seg cells per
 * thread, *_lines lines per block (segments*lines <= 512).  No upstream
 * counterpart. */
int hs2_plan_set_config(hs2_plan *plan, int x_cpl, int x_lanes, int x_lines, int x_lines_y,
                        int y_seg, int y_lines, int z_seg, int z_lines);

/* Plan options (no upstream counterpart).
 *   HS2_OPT_COMBINED_X (default 1): with class maps, let the x sweep of
 *     source-free steps read one combined class byte per cell (the id of
 *     the cell's z, y, x class triple, built by hs2_plan_set_compact when
 *     there are 256 triples or fewer) instead of the three class planes;
 *     0 reads the three planes.  Same numbers either way. */
enum { HS2_OPT_COMBINED_X = 1 };
int hs2_plan_set_option(hs2_plan *plan, int option, int value);

/* info[0..13] = mode (0 none, 1 class maps, 2 dense, 3 dense per side),
 * classes per axis z/y/x, x_cpl, x_lanes, y_seg, y_lines, z_seg, z_lines,
 * x_lines, x_lines_y, 1 when steps without sources hand the y sweep its
 * right-hand side M T* - (Ly/2) T^n from the x sweep (16 B per cell less field
 * traffic per step than writing T* and re-reading T^n) else 0, and the
 * number of combined class ids the x sweep of source-free steps reads (0:
 * it reads the three class planes). */
int hs2_plan_info(const hs2_plan *plan, int *info);

/* One full Douglas step T^n -> T^{n+1} (x, y, z substeps).
 *   d_Tin, d_Tout: n doubles (may alias: in-place is supported);
 *   source (optional, the volumetric power density at this step's time):
 *     S_c = vtab[vmap[c]] + sfield[c] for COMPUTE cells;
 *     d_vmap: uint8 class map or NULL; h_vtab: nvtab host doubles (nvtab<=256);
 *     d_sfield: n doubles or NULL;
 *   d_surf: optional (ny,nx) output of the insulating-top surface readout of
 *     T^{n+1}, fused into the z substep.
 * Replaces run_adi_steps (alternatingdirection_c_pyx.pyx:287-416). */
int hs2_step(hs2_plan *plan, const double *d_Tin, double *d_Tout, const uint8_t *d_vmap,
             const double *h_vtab, int nvtab, const double *d_sfield, double *d_surf,
             void *stream);

/* nsteps steps in place on d_T.  h_vtab_steps (nsteps x nvtab, host) gives
 * each step's per-class source values, or NULL for no source; steps whose
 * row is all zero run without reading the class map.  d_surf_hist
 * (nsteps x ny x nx) receives the readout after every step, or NULL.
 * Replaces the user-side loop over run_adi_steps
 * (demos/testheatsim2.py:105-110). */
int hs2_run(hs2_plan *plan, double *d_T, int nsteps, const uint8_t *d_vmap,
            const double *h_vtab_steps, int nvtab, double *d_surf_hist, void *stream);

/* Device time of the substep kernels, in milliseconds: nsteps source-free
 * steps in place on d_T, like hs2_run, timed with HIP events on `stream`.
 *   method 0: an event pair around every substep kernel (events with the
 *             default system-scope fence, whose cache writeback and
 *             invalidation lengthen the kernel each pair brackets);
 *   method 1: the same with hipEventDisableSystemFence events;
 *   method 2: four runs of nsteps steps, each between one event pair: plain,
 *             and with the x, y or z kernel launched a second time per step
 *             (into a scratch field the call allocates); a kernel's time is
 *             its run's excess over the plain run per step.
 * kernel_ms[0..2] receive the mean time per launch of the x, y and z sweep,
 * kernel_ms[3] the mean time per step (plain run, or first to last event).
 * The field advances 4*nsteps steps under method 2, nsteps otherwise.  A
 * measurement entry point with no upstream counterpart (the upstream has no
 * timers). */
int hs2_run_timed(hs2_plan *plan, double *d_T, int nsteps, int method, double *kernel_ms,
                  void *stream);

/* ---- Gradients (no upstream counterpart: heatsim2 has no adjoint).
 * The differentiable time loop runs nsteps source-driven steps
 *   T^{n+1} = step(T^n; M, g, S^n),   hist[n] = readout(T^{n+1})
 * (readout: the insulating-top surface temperature of
 * surface_temperature.py:31-36), where S^n is the source field d_src[s]
 * (n doubles, W m^-3) of the source s with h_src_step[s] == n (at most one
 * per step), or none.  Given dJ/dhist, hs2_grad_backward returns dJ/dM,
 * dJ/dg per axis (for each cell's minus face, as in hs2_plan_set_dense),
 * dJ/dS for every source and dJ/dT^0, through a hand-written adjoint step:
 * transposed line solves in reverse axis order plus the adjoint stencil,
 * with the forward states recomputed from checkpoints taken every
 * ckpt_every steps.  Works in both coefficient modes; gradients with
 * respect to per-voxel parameters need hs2_plan_set_dense. */

/* Doubles of device workspace hs2_grad_backward needs: (3*ckpt_every+6)*n
 * (each segment's states T^n, T*, T** are recomputed once and kept). */
size_t hs2_grad_work_doubles(const hs2_plan *plan, int ckpt_every);

/* Forward pass: nsteps steps in place on d_T (T^0 in, T^nsteps out),
 * writing T^{k*ckpt_every} to d_ckpt + k*n (ceil(nsteps/ckpt_every) fields)
 * and, if d_surf_hist is not NULL, the readout after every step
 * (nsteps x ny x nx). */
int hs2_grad_forward(hs2_plan *plan, double *d_T, int nsteps, int ckpt_every, double *d_ckpt,
                     const double *d_src, const int32_t *h_src_step, int nsrc,
                     double *d_surf_hist, void *stream);

/* Backward pass over the run that hs2_grad_forward recorded in d_ckpt (same
 * plan coefficients, nsteps, ckpt_every and sources).  d_grad_hist: dJ/dhist
 * (nsteps x ny x nx).  Writes d_gM, d_gz, d_gy, d_gx (n each), d_gsrc
 * (nsrc x n; may be NULL when nsrc == 0) and, if not NULL, d_gT0 (n).
 * d_work: hs2_grad_work_doubles() doubles of scratch.  Needs nz >= 2 and
 * every axis length <= 1024. */
int hs2_grad_backward(hs2_plan *plan, int nsteps, int ckpt_every, const double *d_ckpt,
                      const double *d_src, const int32_t *h_src_step, int nsrc,
                      const double *d_grad_hist, double *d_gM, double *d_gz, double *d_gy,
                      double *d_gx, double *d_gsrc, double *d_gT0, double *d_work,
                      size_t work_doubles, void *stream);

/* ---- Batches (no upstream counterpart).  Many independent problems of
 * different sizes and step counts, each with per-voxel coefficients, run
 * forward and adjoint together: one launch per sweep covers every
 * instance, and the time loop is a captured HIP graph (a chunk of forward
 * steps, or one checkpoint segment of the backward pass) replayed on an
 * internal stream ordered after `stream`.  Axis lengths 1..128 (nz >= 2).
 * All fields are the instances' fields concatenated at the offsets that
 * hs2_batch_layout reports (16-byte aligned; padding cells are ignored). */
typedef struct hs2_batch hs2_batch;

typedef struct {
  int32_t nz, ny, nx;     /* grid                                              */
  int32_t nsteps;         /* time steps of this instance                       */
  int32_t src_step;       /* step whose source field applies, or -1            */
  double dz;              /* z spacing (readout)                               */
} hs2_batch_inst;

/* ckpt_every: steps per checkpoint segment (rounded up to even). */
int hs2_batch_create(const hs2_batch_inst *inst, int ninst, int ckpt_every, hs2_batch **out);
void hs2_batch_destroy(hs2_batch *batch);

/* cell_off[ninst+1], hist_off[ninst+1] (may be NULL): first cell of each
 * instance in the concatenated fields and first value of its surface
 * history (nsteps x ny x nx); the last entries are the totals.  sizes[5]
 * (may be NULL): total cells, total history values, work doubles,
 * checkpoint interval, largest step count. */
int hs2_batch_layout(const hs2_batch *batch, int64_t *cell_off, int64_t *hist_off,
                     int64_t *sizes);

/* Concatenated per-voxel coefficients, as hs2_plan_set_dense. */
int hs2_batch_set_dense(hs2_batch *batch, const double *d_M, const uint8_t *d_w,
                        const double *d_gz, const double *d_gy, const double *d_gx);

/* use_graph 0: launch every step from the host instead; chunk: forward
 * steps per captured graph. */
int hs2_batch_set_options(hs2_batch *batch, int use_graph, int chunk);

/* Forward: d_T (concatenated T^0 in, final fields out), d_src (concatenated
 * source fields, or NULL), d_hist (concatenated surface histories).  d_work
 * keeps the checkpoints for hs2_batch_backward. */
int hs2_batch_forward(hs2_batch *batch, double *d_T, const double *d_src, double *d_hist,
                      double *d_work, size_t work_doubles, void *stream);

/* Backward over the last forward (same coefficients, sources and work):
 * writes dJ/dM, dJ/dg per axis and dJ/dS (d_gsrc may be NULL), all
 * concatenated, from dJ/dhist. */
int hs2_batch_backward(hs2_batch *batch, const double *d_src, const double *d_grad_hist,
                       double *d_gM, double *d_gz, double *d_gy, double *d_gx, double *d_gsrc,
                       double *d_work, size_t work_doubles, void *stream);

/* One Douglas step on per-voxel coefficients that allocates nothing: no
 * plan, no hipMalloc or hipFree, no allocating library routine.  It has no
 * counterpart in the upstream; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Same numbers and kernels as hs2_step on a plan given these coefficients
 * with hs2_plan_set_dense (d_M, d_w, d_gz, d_gy, d_gx as there), with a
 * source field d_sfield (n doubles, W m^-3, or NULL) and the readout d_surf
 * ((ny,nx), or NULL).  d_Tin may equal d_Tout.  The workspace is a fixed
 * device array of 2^21 cells (16 MiB) that each GPU the library runs on
 * holds until the process ends: grids above 2^21 cells are refused with
 * HS2_E_UNSUPPORTED, and calls that overlap on one GPU share the array and
 * race (serialise them, e.g. on one stream).  grid->dt is not used. */
int hs2_step_noalloc(const hs2_grid_desc *grid, const double *d_M, const uint8_t *d_w,
                     const double *d_gz, const double *d_gy, const double *d_gx,
                     const double *d_Tin, double *d_Tout, const double *d_sfield, double *d_surf,
                     void *stream);

/* Insulating-top surface readout of a (nz,ny,nx) field into (ny,nx):
 * T0 - 5/16 (T1 - T0).  Replaces insulating_z_min_surface_temperature
 * (surface_temperature.py:4-36). */
int hs2_surface_temperature(const double *d_T, int nz, int ny, int nx, double dz,
                            double *d_out, void *stream);

/* Message for the last failed call on this thread ("" if none). */
const char *hs2_last_error(void);

/* Library version string. */
const char *hs2_version(void);

#ifdef __cplusplus
}
#endif
#endif
