/* CPU restatement of heatsim2's 3-D ADI Crank-Nicolson step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), and the cpu_baseline leg
 * of bench.py).  Nothing in the product (heatsim2_amd/) links, loads or calls
 * it.
 *
 * Per-cell coefficient record (8 doubles, row-major n x 8), as seen from the
 * cell's own equation:
 *   [0] M    = rho*c/dt (COMPUTE) or 1 (FIXED)        crank_nicolson.pyx:378,481
 *   [1] D    = coefficient of volumetric_source: 1 (COMPUTE) or 0 (FIXED)
 *   [2] gzm  [3] gzp  [4] gym  [5] gyp  [6] gxm  [7] gxp
 *        face conductance per volume (W m^-3 K^-1) of the cell's minus / plus
 *        face on each axis, so that the cell's heat-flow expression is
 *        sum_faces g*(T_neighbour - T_cell) + S           crank_nicolson.pyx:359-363
 */
#ifndef HS2_ORACLE_H
#define HS2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BND_CONDUCTING = 0, ORC_BND_ANISO = 1, ORC_BND_INSULATING = 2, ORC_BND_THIN = 3 };

enum {
  ORC_OK = 0,
  ORC_E_MATERIAL_INDEX = 1,   /* material map entry >= nmat                     */
  ORC_E_BOUNDARY_INDEX = 2,   /* face map entry >= nb                           */
  ORC_E_K_TYPE = 3,           /* tensor k on an isotropic face law, or scalar k
                                 on the anisotropic law (expression.py:311,737) */
  ORC_E_OUTER_FACE = 4,       /* non-zero flux through an outer face of a
                                 COMPUTE cell (alternatingdirection_c.c:160)    */
  ORC_E_ALLOC = 5,
  ORC_E_ARG = 6
};

typedef struct {
  int fixed;       /* TEMPERATURE_FIXED (1) or TEMPERATURE_COMPUTE (0)          */
  int tensor;      /* k given as a diagonal 3x3 matrix (1) or a scalar (0)      */
  double k[3];     /* (kz, ky, kx); all equal for a scalar k                    */
  double rho, c;
} orc_material;

typedef struct {
  int kind;        /* ORC_BND_*                                                 */
  double h;        /* thin-layer conductance W m^-2 K^-1 (ORC_BND_THIN only)    */
} orc_boundary;

int orc_assemble_cells(int nz, int ny, int nx, double dz, double dy, double dx, double dt,
                       const orc_material *mats, int nmat, const orc_boundary *bnds, int nb,
                       const uint8_t *mat, const uint8_t *bz, const uint8_t *by,
                       const uint8_t *bx, double *coef);

/* Reference-structured step: assembled matrices in each substep's permuted
 * row order, CSR sparse products, one global Thomas solve with a precomputed
 * LU, and the transposes -- the structure of
 * alternatingdirection_c_pyx.pyx:389-416 / tridiag.pyx:10-69. */
typedef struct orc_refplan orc_refplan;
orc_refplan *orc_ref_create(int nz, int ny, int nx, const double *coef);
void orc_ref_destroy(orc_refplan *p);
/* Tin, S (volumetric power per cell, W/m^3, may be NULL), Tout: n doubles. */
void orc_ref_step(orc_refplan *p, const double *Tin, const double *S, double *Tout);
/* Export substep s's operators: A (n x 3), D (n), and B / C[c] in CSR
 * (duplicates summed).  nnz_out receives the entry counts; pass NULL arrays
 * to query sizes first. which: 0 = B, 1 = C[0], 2 = C[1]. */
int orc_ref_export_A(orc_refplan *p, int s, double *A, double *D);
int64_t orc_ref_export_csr(orc_refplan *p, int s, int which, int64_t *indptr,
                           int64_t *indices, double *data);

/* Delta-form step (algebraically identical split, per-line Thomas), OpenMP
 * over lines.  work: 2*n doubles. nthreads <= 0 uses the OpenMP default. */
void orc_fast_step(int nz, int ny, int nx, const double *coef, const double *Tin,
                   const double *S, double *Tout, double *work, int nthreads);
/* nsteps of orc_fast_step with an optional per-step source (S_steps is
 * nsteps x n or NULL) and the insulating-top surface readout history
 * (nsteps x ny x nx, may be NULL). */
void orc_fast_run(int nz, int ny, int nx, const double *coef, double *T, int nsteps,
                  const double *S_first, int src_steps, double dz, double *surf_hist,
                  int nthreads);

/* Thomas solve of one tridiagonal system (tridiag.pyx:10-69 restated):
 * rows a[i]*x[i-1] + b[i]*x[i] + c[i]*x[i+1] = r[i]. */
void orc_thomas(int n, const double *a, const double *b, const double *c, const double *r,
                double *x);

#ifdef __cplusplus
}
#endif
#endif
