/* CPU restatement of heatsim2's 3-D ADI Crank-Nicolson step (the checker).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py to check and time the HIP path; the product
 * never links or calls it.  Every function names the upstream lines it
 * restates (paths relative to the heatsim2 source tree).
 *
 * Parity pins (see tests/test_oracle.py): the assembled operators against the
 * upstream's own Amat/Bmat/Cmats/Dvec fixtures, the step against upstream
 * run_adi_steps trajectories, the Thomas solve against dense solves, and the
 * step against energy conservation, constant-field invariance and the exact
 * 1-D Crank-Nicolson reduction.
 */
#include "hs2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(k, j, i) (((size_t)(k) * (size_t)ny + (size_t)(j)) * (size_t)nx + (size_t)(i))

/* Face conductance per volume of one face, from one cell's point of view.
 * axis: 0=z, 1=y, 2=x.  km/kp: the minus- and plus-side materials after the
 * neighbour fallback.  Restates the four face laws:
 *   boundary_conducting.py:18-26        q = -(km+kp)/2 * (Tp-Tm)/d
 *   boundary_conducting_anisotropic.py:22-24,42-44,63-65  (K[a,a] terms)
 *   boundary_insulating.py:2-34         q = 0
 *   boundary_thininsulatinglayer.py:18-23  q = -h (Tp-Tm)
 * each then divided by the spacing in the heat-flow balance
 * (crank_nicolson.pyx:359-363), giving g such that the cell's balance holds
 * g*(T_nb - T_cell). */
static int face_g(const orc_boundary *b, int axis, const orc_material *km,
                  const orc_material *kp, double d, double *g) {
  switch (b->kind) {
    case ORC_BND_INSULATING:
      *g = 0.0;
      return ORC_OK;
    case ORC_BND_THIN:
      *g = b->h * (1.0 / d);
      return ORC_OK;
    case ORC_BND_CONDUCTING:
      if (km->tensor || kp->tensor) return ORC_E_K_TYPE;
      *g = (km->k[axis] + kp->k[axis]) * 0.5 * (1.0 / d) * (1.0 / d);
      return ORC_OK;
    case ORC_BND_ANISO:
      if (!km->tensor || !kp->tensor) return ORC_E_K_TYPE;
      *g = (km->k[axis] + kp->k[axis]) * 0.5 * (1.0 / d) * (1.0 / d);
      return ORC_OK;
    default:
      return ORC_E_ARG;
  }
}

/* Per-cell equation coefficients: the per-cell loop of
 * crank_nicolson.pyx:272-491.  COMPUTE cells (:279-471): the six face
 * classes (:288-296), the neighbour-k fallback -- a neighbour's k is used
 * when it is in range and COMPUTE, otherwise the cell's own k (:298-338) --,
 * the heat-flow balance (:359-363) and the time term rho*c/dt (:378).
 * FIXED cells (:473-486): zero heat flow and time term -(T+ - T-), i.e.
 * M = 1, no couplings and no source. */
int orc_assemble_cells(int nz, int ny, int nx, double dz, double dy, double dx, double dt,
                       const orc_material *mats, int nmat, const orc_boundary *bnds, int nb,
                       const uint8_t *mat, const uint8_t *bz, const uint8_t *by,
                       const uint8_t *bx, double *coef) {
  if (nz <= 0 || ny <= 0 || nx <= 0 || !mats || !bnds || !coef) return ORC_E_ARG;
  const double dd[3] = {dz, dy, dx};
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t c = IDX(k, j, i);
        double *q = coef + 8 * c;
        if (mat[c] >= nmat) return ORC_E_MATERIAL_INDEX;
        const orc_material *me = &mats[mat[c]];
        if (me->fixed) {
          q[0] = 1.0;
          for (int t = 1; t < 8; t++) q[t] = 0.0;
          continue;
        }
        /* neighbours: (minus, plus) per axis z, y, x */
        int inb[3][2] = {{k > 0, k < nz - 1}, {j > 0, j < ny - 1}, {i > 0, i < nx - 1}};
        size_t nbc[3][2] = {{inb[0][0] ? IDX(k - 1, j, i) : 0, inb[0][1] ? IDX(k + 1, j, i) : 0},
                            {inb[1][0] ? IDX(k, j - 1, i) : 0, inb[1][1] ? IDX(k, j + 1, i) : 0},
                            {inb[2][0] ? IDX(k, j, i - 1) : 0, inb[2][1] ? IDX(k, j, i + 1) : 0}};
        /* face class indices: bz[k], bz[k+1], by[j], by[j+1], bx[i], bx[i+1] */
        size_t fz = ((size_t)k * ny + j) * nx + i;
        size_t fy = ((size_t)k * (ny + 1) + j) * nx + i;
        size_t fx = ((size_t)k * ny + j) * (nx + 1) + i;
        int bcls[3][2] = {{bz[fz], bz[fz + (size_t)ny * nx]},
                          {by[fy], by[fy + nx]},
                          {bx[fx], bx[fx + 1]}};
        q[0] = me->rho * me->c * (1.0 / dt);
        q[1] = 1.0;
        for (int a = 0; a < 3; a++)
          for (int s = 0; s < 2; s++) {
            if (bcls[a][s] >= nb) return ORC_E_BOUNDARY_INDEX;
            const orc_material *nbm = me;
            if (inb[a][s]) {
              int mi = mat[nbc[a][s]];
              if (mi >= nmat) return ORC_E_MATERIAL_INDEX;
              if (!mats[mi].fixed) nbm = &mats[mi];
            }
            const orc_material *km = s == 0 ? nbm : me;
            const orc_material *kp = s == 0 ? me : nbm;
            double g;
            int rc = face_g(&bnds[bcls[a][s]], a, km, kp, dd[a], &g);
            if (rc) return rc;
            if (!inb[a][s] && g != 0.0) return ORC_E_OUTER_FACE;
            q[2 + 2 * a + s] = g;
          }
      }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Reference-structured plan: alternatingdirection_c.c + scipy CSR +
 * tridiag.pyx.                                                        */

typedef struct {
  int64_t *indptr, *indices;
  double *data;
  int64_t nnz;
} csr_t;

typedef struct {
  double *A;       /* n x 3, permuted row order           */
  double *L, *U;   /* tridiaglu factors, n x 3 each        */
  double *D;       /* n                                    */
  csr_t B, C[2];
  int nC;
  int perm[3];     /* permuteorder (alternatingdirection_c_pyx.pyx:426-432) */
  int pshape[3];
} orc_sub;

struct orc_refplan {
  int nz, ny, nx;
  size_t n;
  orc_sub sub[3];
  double *tv, *rhs, *res[3], *tmp;
};

/* permuted row index of cell (k,j,i): alternatingdirection_c.c:114 */
static size_t prow(const orc_sub *s, const int64_t pos[3]) {
  return ((size_t)s->pshape[1] * (size_t)pos[s->perm[0]] + (size_t)pos[s->perm[1]]) *
             (size_t)s->pshape[2] +
         (size_t)pos[s->perm[2]];
}

/* Rows with at most `cap` entries each, filled in any order, then sorted by
 * column with duplicates summed -- what coo_matrix(...).tocsr() gives
 * (alternatingdirection_c_pyx.pyx:251,269). */
typedef struct {
  int64_t *cols;
  double *vals;
  uint8_t *cnt;
  int cap;
} rowbuf_t;

static int rowbuf_init(rowbuf_t *b, size_t n, int cap) {
  b->cap = cap;
  b->cols = malloc(sizeof(int64_t) * n * (size_t)cap);
  b->vals = malloc(sizeof(double) * n * (size_t)cap);
  b->cnt = calloc(n, 1);
  return (b->cols && b->vals && b->cnt) ? ORC_OK : ORC_E_ALLOC;
}

static void rowbuf_free(rowbuf_t *b) {
  free(b->cols); free(b->vals); free(b->cnt);
}

static void rowbuf_add(rowbuf_t *b, size_t r, int64_t col, double v) {
  size_t at = r * (size_t)b->cap + b->cnt[r]++;
  b->cols[at] = col;
  b->vals[at] = v;
}

static int rowbuf_to_csr(rowbuf_t *b, size_t n, csr_t *out) {
  out->indptr = malloc(sizeof(int64_t) * (n + 1));
  size_t total = 0;
  for (size_t r = 0; r < n; r++) total += b->cnt[r];
  out->indices = malloc(sizeof(int64_t) * (total ? total : 1));
  out->data = malloc(sizeof(double) * (total ? total : 1));
  if (!out->indptr || !out->indices || !out->data) return ORC_E_ALLOC;
  int64_t w = 0;
  out->indptr[0] = 0;
  for (size_t r = 0; r < n; r++) {
    int64_t *c = b->cols + r * (size_t)b->cap;
    double *v = b->vals + r * (size_t)b->cap;
    int m = b->cnt[r];
    for (int x = 1; x < m; x++) { /* insertion sort, stable */
      int64_t cc = c[x];
      double vv = v[x];
      int y = x - 1;
      while (y >= 0 && c[y] > cc) { c[y + 1] = c[y]; v[y + 1] = v[y]; y--; }
      c[y + 1] = cc;
      v[y + 1] = vv;
    }
    for (int x = 0; x < m; x++) {
      if (x > 0 && c[x] == c[x - 1]) {
        out->data[w - 1] += v[x];
      } else {
        out->indices[w] = c[x];
        out->data[w] = v[x];
        w++;
      }
    }
    out->indptr[r + 1] = w;
  }
  out->nnz = w;
  return ORC_OK;
}

/* tridiag.pyx:10-43 */
static void tridiaglu(size_t n, const double *A, double *L, double *U) {
  memcpy(U, A, sizeof(double) * 3 * n);
  memset(L, 0, sizeof(double) * 3 * n);
  for (size_t row = 0; row < n; row++) {
    L[3 * row + 1] = U[3 * row + 1];
    U[3 * row + 2] /= U[3 * row + 1];
    U[3 * row + 1] = 1.0;
    if (row < n - 1) {
      double coefficient = U[3 * (row + 1) + 0];
      U[3 * (row + 1) + 0] = 0.0;
      U[3 * (row + 1) + 1] -= U[3 * row + 2] * coefficient;
      L[3 * (row + 1) + 0] = coefficient;
    }
  }
}

/* tridiag.pyx:46-69 */
static void tridiagsolve(size_t n, const double *L, const double *U, const double *b,
                         double *x) {
  x[0] = b[0] / L[1];
  for (size_t row = 1; row < n; row++) x[row] = (b[row] - L[3 * row] * x[row - 1]) / L[3 * row + 1];
  for (size_t row = n - 1; row-- > 0;) x[row] = x[row] - U[3 * row + 2] * x[row + 1];
}

static void ref_free_sub(orc_sub *s) {
  free(s->A); free(s->L); free(s->U); free(s->D);
  free(s->B.indptr); free(s->B.indices); free(s->B.data);
  for (int c = 0; c < 2; c++) { free(s->C[c].indptr); free(s->C[c].indices); free(s->C[c].data); }
}

void orc_ref_destroy(orc_refplan *p) {
  if (!p) return;
  for (int s = 0; s < 3; s++) ref_free_sub(&p->sub[s]);
  free(p->tv); free(p->rhs); free(p->tmp);
  for (int s = 0; s < 3; s++) free(p->res[s]);
  free(p);
}

/* Build the three substeps' A, B, C[], D from per-cell coefficients, with the
 * entry conventions of alternatingdirection_c.c:102-199 applied to the
 * Douglas split of alternatingdirection_c_pyx.pyx:475-493: substep s solves
 * axis x, y, z for s = 0, 1, 2; axes solved in an earlier substep u < s
 * contribute 1/2 (p<u> -> C[u]) + 1/2 (m -> B); the axis solved now gives
 * 1/2 (p<s> -> A, sign flipped) + 1/2 (m -> B); later axes are explicit
 * (bare names -> B, full weight); the time term gives M at p<s> (A) and at m
 * (B); the volumetric_source coefficient is D. */
orc_refplan *orc_ref_create(int nz, int ny, int nx, const double *coef) {
  orc_refplan *p = calloc(1, sizeof(*p));
  if (!p) return NULL;
  p->nz = nz; p->ny = ny; p->nx = nx;
  size_t n = (size_t)nz * ny * nx;
  p->n = n;
  const int perms[3][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}};
  const int shape[3] = {nz, ny, nx};
  /* coefficient slots for axis order x (substep 0), y (1), z (2) */
  const int gm_slot[3] = {6, 4, 2}, gp_slot[3] = {7, 5, 3};
  const int axis_of[3] = {2, 1, 0};  /* substep -> array axis (z=0,y=1,x=2) */
  for (int s = 0; s < 3; s++) {
    orc_sub *S = &p->sub[s];
    for (int t = 0; t < 3; t++) { S->perm[t] = perms[s][t]; S->pshape[t] = shape[perms[s][t]]; }
    S->nC = s;
    S->A = calloc(3 * n, sizeof(double));
    S->L = malloc(3 * n * sizeof(double));
    S->U = malloc(3 * n * sizeof(double));
    S->D = calloc(n, sizeof(double));
    rowbuf_t eB, eC[2];
    int rc = rowbuf_init(&eB, n, 10);
    for (int c = 0; c < s && rc == ORC_OK; c++) rc = rowbuf_init(&eC[c], n, 4);
    if (rc != ORC_OK || !S->A || !S->L || !S->U || !S->D) {
      rowbuf_free(&eB);
      for (int c = 0; c < s; c++) rowbuf_free(&eC[c]);
      orc_ref_destroy(p);
      return NULL;
    }
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const double *q = coef + 8 * IDX(k, j, i);
          int64_t pos[3] = {k, j, i};
          size_t r = prow(S, pos);
          if (q[1] == 0.0) { /* FIXED: A diag 1, B self +1 (crank_nicolson.pyx:481) */
            S->A[3 * r + 1] = 1.0;
            rowbuf_add(&eB, r, (int64_t)r, 1.0);
            continue;
          }
          double self_m = q[0], self_bare = 0.0;
          double adiag = q[0];
          for (int u = 0; u < 3; u++) { /* u = substep that solves this axis */
            double gm = q[gm_slot[u]], gp = q[gp_slot[u]];
            int ax = axis_of[u];
            for (int side = 0; side < 2; side++) {
              double g = side ? gp : gm;
              if (g == 0.0) continue;
              int64_t np_[3] = {k, j, i};
              np_[ax] += side ? 1 : -1;
              size_t col = prow(S, np_);
              if (u < s) {
                rowbuf_add(&eC[u], r, (int64_t)col, 0.5 * g);
                rowbuf_add(&eC[u], r, (int64_t)r, -0.5 * g);
                rowbuf_add(&eB, r, (int64_t)col, 0.5 * g);
                self_m -= 0.5 * g;
              } else if (u == s) {
                S->A[3 * r + (side ? 2 : 0)] = -0.5 * g;
                adiag += 0.5 * g;
                rowbuf_add(&eB, r, (int64_t)col, 0.5 * g);
                self_m -= 0.5 * g;
              } else {
                rowbuf_add(&eB, r, (int64_t)col, g);
                self_bare -= g;
              }
            }
          }
          S->A[3 * r + 1] = adiag;
          rowbuf_add(&eB, r, (int64_t)r, self_m);
          if (self_bare != 0.0) rowbuf_add(&eB, r, (int64_t)r, self_bare);
          S->D[r] = q[1];
        }
    rc = rowbuf_to_csr(&eB, n, &S->B);
    for (int c = 0; c < s && rc == ORC_OK; c++) rc = rowbuf_to_csr(&eC[c], n, &S->C[c]);
    rowbuf_free(&eB);
    for (int c = 0; c < s; c++) rowbuf_free(&eC[c]);
    if (rc) { orc_ref_destroy(p); return NULL; }
    tridiaglu(n, S->A, S->L, S->U);
  }
  p->tv = malloc(n * sizeof(double));
  p->rhs = malloc(n * sizeof(double));
  p->tmp = malloc(n * sizeof(double));
  for (int s = 0; s < 3; s++) p->res[s] = malloc(n * sizeof(double));
  if (!p->tv || !p->rhs || !p->tmp || !p->res[0] || !p->res[1] || !p->res[2]) {
    orc_ref_destroy(p);
    return NULL;
  }
  return p;
}

/* array.transpose(perm).reshape(n) (alternatingdirection_c_pyx.pyx:397) */
static void permute_to(const orc_refplan *p, const orc_sub *S, const double *src, double *dst) {
  int nz = p->nz, ny = p->ny, nx = p->nx;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int64_t pos[3] = {k, j, i};
        dst[prow(S, pos)] = src[IDX(k, j, i)];
      }
}

static void permute_from(const orc_refplan *p, const orc_sub *S, const double *src, double *dst) {
  int nz = p->nz, ny = p->ny, nx = p->nx;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int64_t pos[3] = {k, j, i};
        dst[IDX(k, j, i)] = src[prow(S, pos)];
      }
}

static void csr_matvec_add(size_t n, const csr_t *M, const double *x, double *y, int accumulate) {
  for (size_t r = 0; r < n; r++) {
    double acc = accumulate ? y[r] : 0.0;
    double s = 0.0;
    for (int64_t t = M->indptr[r]; t < M->indptr[r + 1]; t++) s += M->data[t] * x[M->indices[t]];
    y[r] = acc + s;
  }
}

/* run_adi_steps substep loop: alternatingdirection_c_pyx.pyx:388-416 */
void orc_ref_step(orc_refplan *p, const double *Tin, const double *S, double *Tout) {
  size_t n = p->n;
  for (int s = 0; s < 3; s++) {
    orc_sub *Sb = &p->sub[s];
    permute_to(p, Sb, Tin, p->tv);
    csr_matvec_add(n, &Sb->B, p->tv, p->rhs, 0);
    for (int c = 0; c < s; c++) {
      permute_to(p, Sb, p->res[c], p->tmp);
      csr_matvec_add(n, &Sb->C[c], p->tmp, p->rhs, 1);
    }
    if (S) {
      permute_to(p, Sb, S, p->tmp);
      for (size_t r = 0; r < n; r++) p->rhs[r] += Sb->D[r] * p->tmp[r];
    }
    tridiagsolve(n, Sb->L, Sb->U, p->rhs, p->tmp);
    permute_from(p, Sb, p->tmp, p->res[s]);
  }
  memcpy(Tout, p->res[2], n * sizeof(double));
}

int orc_ref_export_A(orc_refplan *p, int s, double *A, double *D) {
  if (s < 0 || s > 2) return ORC_E_ARG;
  memcpy(A, p->sub[s].A, 3 * p->n * sizeof(double));
  memcpy(D, p->sub[s].D, p->n * sizeof(double));
  return ORC_OK;
}

int64_t orc_ref_export_csr(orc_refplan *p, int s, int which, int64_t *indptr, int64_t *indices,
                           double *data) {
  if (s < 0 || s > 2 || which < 0 || which > s) return -1;
  csr_t *M = which == 0 ? &p->sub[s].B : &p->sub[s].C[which - 1];
  if (indptr) memcpy(indptr, M->indptr, (p->n + 1) * sizeof(int64_t));
  if (indices) memcpy(indices, M->indices, (size_t)M->nnz * sizeof(int64_t));
  if (data) memcpy(data, M->data, (size_t)M->nnz * sizeof(double));
  return M->nnz;
}

/* ------------------------------------------------------------------ */
/* Delta-form step.  Subtracting consecutive Douglas equations
 * (alternatingdirection_c_pyx.pyx:483-493 with expression.py:816-819's
 * 1/2-1/2 weights) gives
 *   (M - Lx/2) T*   = (M + Lx/2 + Ly + Lz) T^n + D S
 *   (M - Ly/2) T**  = M T* - (Ly/2) T^n
 *   (M - Lz/2) T^n+1 = M T** - (Lz/2) T^n
 * where (La w)_c = g_a-(w_{c-e_a} - w_c) + g_a+(w_{c+e_a} - w_c). */

/* Thomas (tridiag.pyx:10-69 restated for one line, LU folded into the sweep). */
void orc_thomas(int n, const double *a, const double *b, const double *c, const double *r,
                double *x) {
  double *cp = malloc(sizeof(double) * (size_t)n);
  double den = b[0];
  cp[0] = c[0] / den;
  x[0] = r[0] / den;
  for (int i = 1; i < n; i++) {
    den = b[i] - a[i] * cp[i - 1];
    cp[i] = c[i] / den;
    x[i] = (r[i] - a[i] * x[i - 1]) / den;
  }
  for (int i = n - 2; i >= 0; i--) x[i] -= cp[i] * x[i + 1];
  free(cp);
}

/* One line of length L starting at cell `base` with stride `stride`, solve
 * axis slots gms/gps.  rhs[] must hold the right-hand side; the solution is
 * written to out[] along the line. */
static void solve_line(int L, size_t base, size_t stride, const double *coef, int gms, int gps,
                       const double *rhs, double *out, double *cp, double *dp) {
  for (int t = 0; t < L; t++) {
    const double *q = coef + 8 * (base + (size_t)t * stride);
    double gm = q[gms], gp = q[gps];
    double a = -0.5 * gm, b = q[0] + 0.5 * (gm + gp), c = -0.5 * gp;
    double den = t ? b - a * cp[t - 1] : b;
    cp[t] = c / den;
    dp[t] = (t ? rhs[t] - a * dp[t - 1] : rhs[t]) / den;
  }
  double x = dp[L - 1];
  out[base + (size_t)(L - 1) * stride] = x;
  for (int t = L - 2; t >= 0; t--) {
    x = dp[t] - cp[t] * x;
    out[base + (size_t)t * stride] = x;
  }
}

void orc_fast_step(int nz, int ny, int nx, const double *coef, const double *Tin,
                   const double *S, double *Tout, double *work, int nthreads) {
  size_t n = (size_t)nz * ny * nx, sxy = (size_t)ny * nx;
  double *Ts = work, *Tss = work + n;
  int Lmax = nz > ny ? nz : ny;
  if (nx > Lmax) Lmax = nx;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  /* substep 0: x-lines */
#pragma omp parallel
  {
    double *buf = malloc(sizeof(double) * 3 * (size_t)Lmax);
    double *rhs = buf, *cp = buf + Lmax, *dp = buf + 2 * Lmax;
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nz * ny; line++) {
      int k = (int)(line / ny), j = (int)(line % ny);
      size_t base = IDX(k, j, 0);
      for (int i = 0; i < nx; i++) {
        size_t c = base + i;
        const double *q = coef + 8 * c;
        double T = Tin[c];
        double r = q[0] * T;
        if (q[6] != 0.0) r += 0.5 * q[6] * (Tin[c - 1] - T);
        if (q[7] != 0.0) r += 0.5 * q[7] * (Tin[c + 1] - T);
        if (q[4] != 0.0) r += q[4] * (Tin[c - nx] - T);
        if (q[5] != 0.0) r += q[5] * (Tin[c + nx] - T);
        if (q[2] != 0.0) r += q[2] * (Tin[c - sxy] - T);
        if (q[3] != 0.0) r += q[3] * (Tin[c + sxy] - T);
        if (S) r += q[1] * S[c];
        rhs[i] = r;
      }
      solve_line(nx, base, 1, coef, 6, 7, rhs, Ts, cp, dp);
    }
    /* substep 1: y-lines, (M - Ly/2) T** = M T* - (Ly/2) T^n */
#pragma omp for schedule(static)
    for (long line = 0; line < (long)nz * nx; line++) {
      int k = (int)(line / nx), i = (int)(line % nx);
      size_t base = IDX(k, 0, i);
      for (int j = 0; j < ny; j++) {
        size_t c = base + (size_t)j * nx;
        const double *q = coef + 8 * c;
        double T = Tin[c];
        double r = q[0] * Ts[c];
        if (q[4] != 0.0) r -= 0.5 * q[4] * (Tin[c - nx] - T);
        if (q[5] != 0.0) r -= 0.5 * q[5] * (Tin[c + nx] - T);
        rhs[j] = r;
      }
      solve_line(ny, base, nx, coef, 4, 5, rhs, Tss, cp, dp);
    }
    /* substep 2: z-lines, (M - Lz/2) T^n+1 = M T** - (Lz/2) T^n */
#pragma omp for schedule(static)
    for (long line = 0; line < (long)ny * nx; line++) {
      size_t base = (size_t)line;
      for (int k = 0; k < nz; k++) {
        size_t c = base + (size_t)k * sxy;
        const double *q = coef + 8 * c;
        double T = Tin[c];
        double r = q[0] * Tss[c];
        if (q[2] != 0.0) r -= 0.5 * q[2] * (Tin[c - sxy] - T);
        if (q[3] != 0.0) r -= 0.5 * q[3] * (Tin[c + sxy] - T);
        rhs[k] = r;
      }
      solve_line(nz, base, sxy, coef, 2, 3, rhs, Tout, cp, dp);
    }
    free(buf);
  }
}

/* surface_temperature.py:31-36 restated: average of the quadratic (zero
 * slope at z=0) and linear extrapolations from the first two layers. */
static void surface_readout(int ny, int nx, const double *T, double dz, double *out) {
  size_t sxy = (size_t)ny * nx;
  for (size_t c = 0; c < sxy; c++) {
    double T0 = T[c], T1 = T[c + sxy];
    double a = (T1 - T0) / (2.0 * pow(dz, 2.0));
    double cc = T0 - .25 * a * pow(dz, 2.0);
    double dT = T1 - T0;
    double linearext = T0 - dT / 2.0;
    out[c] = (linearext + cc) / 2.0;
  }
}

void orc_fast_run(int nz, int ny, int nx, const double *coef, double *T, int nsteps,
                  const double *S_first, int src_steps, double dz, double *surf_hist,
                  int nthreads) {
  size_t n = (size_t)nz * ny * nx;
  double *work = malloc(sizeof(double) * 3 * n);
  double *Tn = work + 2 * n;
  for (int s = 0; s < nsteps; s++) {
    memcpy(Tn, T, n * sizeof(double));
    orc_fast_step(nz, ny, nx, coef, Tn, s < src_steps ? S_first : NULL, T, work, nthreads);
    if (surf_hist && nz >= 2) surface_readout(ny, nx, T, dz, surf_hist + (size_t)s * ny * nx);
  }
  free(work);
}
