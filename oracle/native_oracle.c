/*
 * CPU restatement of T-Flows' Native_Mod (Sources/Process/Cpu/Native_Mod/).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library and the timed
 * CPU baseline.  Nothing in the product links or calls it.
 *
 * Each routine follows the reference's operation order so that, compiled
 * without floating-point contraction (-ffp-contract=off) and without OpenMP,
 * it performs the same IEEE FP64 operations in the same order as the
 * reference built with its default options (OMP=no, REAL=double,
 * Sources/Process/Cpu/makefile:10-17).  The `omp` pragmas sit where the
 * reference has `!$omp` directives; they only take effect in the separately
 * built OpenMP variant used to time the CPU baseline on all cores.
 *
 * Conventions: C index i is Fortran index i+1.  Index arrays carry 1-based
 * values, so col[j]-1 / row[i]-1 / dia[i]-1 are C positions.
 */
#include "native_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Sources/Shared/Const_Mod.f90:84-93 */
#define OR_HUGE 1.0e+15
#define OR_TINY 1.0e-15
#define OR_MICRO 1.0e-6

/* Summation order of the global reductions.  0 (default) is the reference's
 * serial left-to-right order.  1, 2, 4, 5 (6) sum blocks of 64, 1000, 256,
 * 4096 (1) terms left to right and combine the block sums pairwise; 3 sums
 * serially from the last term to the first.  7 (8) keeps the reference's
 * order but takes the normalisation factor fn one ulp below (above) the
 * reference's: the sum of the diagonal that fn comes from rounds differently
 * in another order, and on ill-conditioned systems one ulp of fn alone moves
 * the iteration count by as much as reordering the dot products.  9 (10) is
 * the fully pairwise order 6 with fn one ulp below (above) its own.  Equally
 * valid answers, used only to measure how far rounding alone moves a solve
 * (the envelope the GPU's own order is compared against).  Not part of the
 * reference. */
static int g_sum_order = 0;

void or_set_sum_order(int order) { g_sum_order = order; }

/* Ranks of a multi-rank run on one global matrix: part[i] is the rank that
 * owns row i (NULL: one rank).  With parts set, "icc" is the reference's
 * per-rank DIC: each rank forms and applies it over its own rows only
 * (Prec_Form.f90:68-78 and Prec_Solve.f90:64-92 on the rank's matrix, whose
 * buffer columns the recurrences never read, :89), which on the global
 * matrix means skipping every coupling between rows of different ranks.
 * Sum order 11 is the reductions of such a run: each rank sums its rows
 * serially (its own loop) and the rank totals are added in rank order (the
 * allreduce); ranks must then own contiguous runs of rows.  Test
 * infrastructure, not part of the reference. */
static const int* g_parts = NULL;

void or_set_parts(const int* part) { g_parts = part; }

static int other_rank(int i, int k) { return g_parts && g_parts[k] != g_parts[i]; }

/* the reductions run in a blocked (not the serial) order */
static int reordered(void) {
  return (g_sum_order >= 1 && g_sum_order <= 6) || g_sum_order == 9 || g_sum_order == 10 ||
         (g_sum_order == 11 && g_parts);
}

/* Progress line on stderr every `every` CG iterations (0: none), so that a
 * long solve (the full-size CPU baseline) reports as it goes.  Not part of
 * the reference. */
static int g_progress = 0;

void or_set_progress(int every) { g_progress = every; }

static double pairwise(const double* s, int n) {
  if (n == 1) return s[0];
  int h = n / 2;
  return pairwise(s, h) + pairwise(s + h, n - h);
}

/* sum_i f(i) in the blocked order (3: serial from the last term to the
 * first); kind: 0 a*b, 1 (a*a)/(b*b), 2 a */
static double blocked(const double* a, const double* b, int n, int kind) {
  if (n <= 0) return 0.0;
  if (g_sum_order == 11) {                      /* per rank, then over ranks */
    double s = 0.0;
    for (int i = 0; i < n;) {
      const int rank = g_parts[i];
      double t = 0.0;
      for (; i < n && g_parts[i] == rank; i++)
        t = t + (kind == 0 ? a[i] * b[i] : (kind == 1 ? (a[i] * a[i]) / (b[i] * b[i]) : a[i]));
      s = s + t;
    }
    return s;
  }
  if (g_sum_order == 3) {
    double s = 0.0;
    for (int i = n - 1; i >= 0; i--)
      s = s + (kind == 0 ? a[i] * b[i] : (kind == 1 ? (a[i] * a[i]) / (b[i] * b[i]) : a[i]));
    return s;
  }
  const int bs = g_sum_order == 2   ? 1000
                 : g_sum_order == 4 ? 256
                 : g_sum_order == 5 ? 4096
                 : g_sum_order == 6 || g_sum_order == 9 || g_sum_order == 10 ? 1
                                    : 64;
  int nb = (n + bs - 1) / bs;
  double* part = (double*)malloc(sizeof(double) * (size_t)nb);
  for (int k = 0; k < nb; k++) {
    double s = 0.0;
    for (int i = bs * k; i < n && i < bs * k + bs; i++) {
      double t = kind == 0 ? a[i] * b[i] : (kind == 1 ? (a[i] * a[i]) / (b[i] * b[i]) : a[i]);
      s = s + t;
    }
    part[k] = s;
  }
  double s = pairwise(part, nb);
  free(part);
  return s;
}

/* Normalized_Root_Mean_Square.f90:2-65 */
double or_nrms(const or_matrix* A, int ni, const double* r, const double* x,
               const double* norm) {
  double rms = 0.0;
  if (reordered()) {
    double* dg = (double*)malloc(sizeof(double) * (size_t)(ni > 0 ? ni : 1));
    for (int i = 0; i < ni; i++) dg[i] = A->val[A->dia[i] - 1];
    rms = blocked(r, dg, ni, 1);
    free(dg);
  } else {
    for (int i = 0; i < ni; i++) {             /* :34-36, serial in the reference */
      double a = A->val[A->dia[i] - 1];
      rms = rms + (r[i] * r[i]) / (a * a);
    }
  }
  rms = sqrt(rms);                              /* :39 */
  double x_min, x_max;
  if (!norm) {                                  /* :42-48 */
    x_min = +OR_HUGE;
    x_max = -OR_HUGE;
    for (int i = 0; i < ni; i++) {
      x_min = fmin(x_min, x[i]);
      x_max = fmax(x_max, x[i]);
    }
  } else {                                      /* :49-52 */
    x_min = 0.0;
    x_max = *norm;
  }
  double x_max_min = x_max - x_min;             /* :59-61 */
  x_max_min = fmax(x_max_min, OR_TINY);
  return rms / x_max_min;
}

/* Residual_Vector.f90:34-42: r_i = b_i - sum_j a_ij x_j, subtracting in CSR
 * order.  (The reference's OpenMP clause leaves k shared, a race; here k is
 * private, which is what the serial reference computes.) */
void or_residual(const or_matrix* A, int ni, double* r, const double* b,
                 const double* x) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    double ri = b[i];
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
      int k = A->col[j] - 1;
      ri = ri - A->val[j] * x[k];
    }
    r[i] = ri;
  }
}

/* Cg.f90:177-185: q_i = 0 + sum_j a_ij p_j in CSR order. */
void or_spmv(const or_matrix* A, int ni, const double* p, double* q) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    double qi = 0.0;
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
      int k = A->col[j] - 1;
      qi = qi + A->val[j] * p[k];
    }
    q[i] = qi;
  }
}

/* Prec_Form.f90:47-89 */
void or_prec_form(const or_matrix* A, int ni, double* d, double* d_inv, int prec) {
  if (prec == OR_PREC_JACOBI) {                 /* :47-53 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) {
      d[i] = A->val[A->dia[i] - 1];
      d_inv[i] = 1.0 / d[i];
    }
  } else if (prec == OR_PREC_ICC) {             /* :68-78, serial recurrence */
    for (int i = 0; i < ni; i++) {
      double sum = A->val[A->dia[i] - 1];
      for (int j = A->row[i]; j < A->row[i + 1] - 1; j++) { /* a_row(i)+1 .. */
        int k = A->col[j] - 1;
        if (k > i) break;
        if (other_rank(i, k)) continue;
        sum = sum - d[k] * A->val[j] * A->val[j];
      }
      d_inv[i] = sum;
      d[i] = 1.0 / sum;
    }
  } else {                                      /* :83-89 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) d[i] = 1.0;
  }
}

/* Prec_Solve.f90:44-103 */
void or_prec_solve(const or_matrix* A, int ni, const double* d,
                   const double* d_inv, double* x, const double* b, int prec) {
  if (prec == OR_PREC_JACOBI) {                 /* :44-49 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) x[i] = b[i] * d_inv[i];
  } else if (prec == OR_PREC_ICC) {             /* :64-92 */
    for (int i = 0; i < ni; i++) {              /* forward substitution */
      double sum = b[i];
      for (int j = A->row[i]; j < A->row[i + 1] - 1; j++) {
        int k = A->col[j] - 1;
        if (k > i) break;
        if (other_rank(i, k)) continue;
        sum = sum - A->val[j] * x[k];
      }
      x[i] = sum * d[i];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) x[i] = x[i] * d_inv[i];
    for (int i = ni - 1; i >= 0; i--) {         /* backward substitution */
      double sum = x[i];
      for (int j = A->row[i + 1] - 2; j >= A->row[i]; j--) { /* down to diag+1 */
        int k = A->col[j] - 1;
        if (k < i) break;
        if (k < ni && !other_rank(i, k)) sum = sum - A->val[j] * x[k];
      }
      x[i] = sum * d[i];
    }
  } else {                                      /* :94-101 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) x[i] = b[i];
  }
}

/* Cg.f90:62-84 (Bicg.f90:63-85): fn = 1/(sum_i a_ii / ni); val *= fn and
 * b *= fn over rows 1..nt. */
static double normalise(or_matrix* A, double* b) {
  double sum_a = 0.0;
  int sum_n = 0;
  if (reordered()) {
    double* dg = (double*)malloc(sizeof(double) * (size_t)A->ni);
    for (int i = 0; i < A->ni; i++) dg[i] = A->val[A->dia[i] - 1];
    sum_a = blocked(dg, NULL, A->ni, 2);
    sum_n = A->ni;
    free(dg);
  } else {
#pragma omp parallel for schedule(static) reduction(+ : sum_a) reduction(+ : sum_n)
    for (int i = 0; i < A->ni; i++) {
      sum_a = sum_a + A->val[A->dia[i] - 1];
      sum_n = sum_n + 1;
    }
  }
  sum_a = sum_a / (double)sum_n;
  double fn = 1.0 / sum_a;
  if (g_sum_order == 7 || g_sum_order == 9) fn = nextafter(fn, 0.0);
  if (g_sum_order == 8 || g_sum_order == 10) fn = nextafter(fn, INFINITY);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A->nt; i++) {
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) A->val[j] = A->val[j] * fn;
    b[i] = b[i] * fn;
  }
  return fn;
}

/* Cg.f90:245-250: de-normalise (serial in the reference's Cg). */
static void denormalise(or_matrix* A, double* b, double fn) {
  for (int i = 0; i < A->nt; i++) {
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) A->val[j] = A->val[j] / fn;
    b[i] = b[i] / fn;
  }
}

static double sumsq(const double* v, int n) {
  if (reordered()) return blocked(v, v, n, 0);
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + v[i] * v[i];
  return s;
}

/* Stopping measure: the reference's NRMS, or ||r||_2/||b||_2 (bnorm given). */
static double measure(const or_matrix* A, int ni, const double* r,
                      const double* x, const double* norm, int crit,
                      double bnorm) {
  if (crit == OR_CRIT_RELRES) return sqrt(sumsq(r, ni)) / bnorm;
  return or_nrms(A, ni, r, x, norm);
}

/* Cg.f90:2-257, with the stopping measure selectable (OR_CRIT_NRMS is the
 * reference's). */
int or_cg_ex(or_matrix* A, double* x, double* b, int prec, int miter,
             int* niter, double tol, double* fin_res, const double* norm,
             int crit) {
  const int nt = A->nt, ni = A->ni;
  /* Work % Connect_Real_Cell zero-fills the scratch arrays (Cg.f90:44) */
  double* p1 = (double*)calloc((size_t)nt, sizeof(double));
  double* q1 = (double*)calloc((size_t)nt, sizeof(double));
  double* r1 = (double*)calloc((size_t)nt, sizeof(double));
  double* d = (double*)calloc((size_t)nt, sizeof(double));
  double* d_inv = (double*)calloc((size_t)nt, sizeof(double));
  if (!p1 || !q1 || !r1 || !d || !d_inv) {
    free(p1); free(q1); free(r1); free(d); free(d_inv);
    return -2;
  }
  double res = 0.0, rho = 0.0, rho_old = 0.0, alfa, beta, bnrm2;
  int iter;

  double fn = normalise(A, b);                  /* :62-84 */
  or_prec_form(A, ni, d, d_inv, prec);          /* :89 */

  double bnorm = sqrt(sumsq(b, ni));
  if (crit == OR_CRIT_RELRES)
    bnrm2 = bnorm > 0.0 ? 1.0 : 0.0;            /* b = 0 solves trivially */
  else
    bnrm2 = or_nrms(A, ni, b, x, norm);         /* :95-99 */
  if (bnrm2 < tol) { iter = 0; goto done; }     /* :101-104 */

  or_residual(A, ni, r1, b, x);                 /* :109 */
  res = measure(A, ni, r1, x, norm, crit, bnorm);   /* :114-118 */
  if (res < tol) { iter = 0; goto done; }       /* :120-123 */

#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) p1[i] = r1[i];   /* :128-132 */

  for (iter = 1; iter <= miter; iter++) {       /* :139 */
    or_prec_solve(A, ni, d, d_inv, q1, r1, prec);   /* :145 */
    rho = 0.0;                                  /* :150-156 */
    if (reordered()) {
      rho = blocked(q1, r1, ni, 0);
    } else {
#pragma omp parallel for schedule(static) reduction(+ : rho)
      for (int i = 0; i < ni; i++) rho = rho + q1[i] * r1[i];
    }
    if (iter == 1) {                            /* :158-163 */
#pragma omp parallel for schedule(static)
      for (int i = 0; i < ni; i++) p1[i] = q1[i];
    } else {                                    /* :164-171 */
      beta = rho / rho_old;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < ni; i++) p1[i] = q1[i] + beta * p1[i];
    }
    or_spmv(A, ni, p1, q1);                     /* :176-185 */
    alfa = 0.0;                                 /* :190-198 */
    if (reordered()) {
      alfa = blocked(p1, q1, ni, 0);
    } else {
#pragma omp parallel for schedule(static) reduction(+ : alfa)
      for (int i = 0; i < ni; i++) alfa = alfa + p1[i] * q1[i];
    }
    alfa = rho / alfa;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ni; i++) {              /* :204-209 */
      x[i] = x[i] + alfa * p1[i];
      r1[i] = r1[i] - alfa * q1[i];
    }
    res = measure(A, ni, r1, x, norm, crit, bnorm);  /* :214-218 */
    if (g_progress > 0 && iter % g_progress == 0) {
      fprintf(stderr, "oracle cg: iteration %d, residual %.3e\n", iter, res);
      fflush(stderr);
    }
    if (res < tol) goto done;                   /* :220 */
    rho_old = rho;                              /* :222 */
  }
  iter = iter - 1;                              /* :228 */

done:
  denormalise(A, b, fn);                        /* :245-250 */
  *fin_res = res;                               /* :252-253 */
  *niter = iter;
  free(p1); free(q1); free(r1); free(d); free(d_inv);
  return 0;
}

int or_cg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
          double tol, double* fin_res, const double* norm) {
  return or_cg_ex(A, x, b, prec, miter, niter, tol, fin_res, norm, OR_CRIT_NRMS);
}

/* Bicg.f90:2-269, restated literally with its shadow vectors. */
int or_bicg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
            double tol, double* fin_res, const double* norm) {
  const int nt = A->nt, ni = A->ni;
  double* p1 = (double*)calloc((size_t)nt, sizeof(double));
  double* q1 = (double*)calloc((size_t)nt, sizeof(double));
  double* r1 = (double*)calloc((size_t)nt, sizeof(double));
  double* p2 = (double*)calloc((size_t)nt, sizeof(double));
  double* q2 = (double*)calloc((size_t)nt, sizeof(double));
  double* r2 = (double*)calloc((size_t)nt, sizeof(double));
  double* d = (double*)calloc((size_t)nt, sizeof(double));
  double* d_inv = (double*)calloc((size_t)nt, sizeof(double));
  if (!p1 || !q1 || !r1 || !p2 || !q2 || !r2 || !d || !d_inv) {
    free(p1); free(q1); free(r1); free(p2); free(q2); free(r2); free(d); free(d_inv);
    return -2;
  }
  double res = 0.0, rho = 0.0, rho_old = 0.0, alfa, beta, bnrm2;
  int iter;

  double fn = normalise(A, b);                  /* :63-85 */
  or_prec_form(A, ni, d, d_inv, prec);          /* :90 */
  bnrm2 = or_nrms(A, ni, b, x, norm);           /* :96-100 */
  if (bnrm2 < tol) { iter = 0; goto done; }
  or_residual(A, ni, r1, b, x);                 /* :110 */
  res = or_nrms(A, ni, r1, x, norm);            /* :115-119 */
  if (res < tol) { iter = 0; goto done; }
  for (int i = 0; i < ni; i++) r2[i] = r1[i];   /* :129-133 */

  for (iter = 1; iter <= miter; iter++) {       /* :140 */
    or_prec_solve(A, ni, d, d_inv, q1, r1, prec);   /* :147 */
    or_prec_solve(A, ni, d, d_inv, q2, r2, prec);   /* :148 */
    rho = 0.0;                                  /* :153-159 */
    for (int i = 0; i < ni; i++) rho = rho + q1[i] * r2[i];
    if (iter == 1) {                            /* :161-167 */
      for (int i = 0; i < ni; i++) { p1[i] = q1[i]; p2[i] = q2[i]; }
    } else {                                    /* :168-176 */
      beta = rho / rho_old;
      for (int i = 0; i < ni; i++) {
        p1[i] = q1[i] + beta * p1[i];
        p2[i] = q2[i] + beta * p2[i];
      }
    }
    for (int i = 0; i < ni; i++) {              /* :184-195, both with A */
      double s1 = 0.0, s2 = 0.0;
      for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
        int k = A->col[j] - 1;
        s1 = s1 + A->val[j] * p1[k];
        s2 = s2 + A->val[j] * p2[k];
      }
      q1[i] = s1;
      q2[i] = s2;
    }
    alfa = 0.0;                                 /* :200-207 */
    for (int i = 0; i < ni; i++) alfa = alfa + p2[i] * q1[i];
    alfa = rho / alfa;
    for (int i = 0; i < ni; i++) {              /* :213-219 */
      x[i] = x[i] + alfa * p1[i];
      r1[i] = r1[i] - alfa * q1[i];
      r2[i] = r2[i] - alfa * q2[i];
    }
    res = or_nrms(A, ni, r1, x, norm);          /* :224-228 */
    if (res < tol) goto done;                   /* :230 */
    rho_old = rho;                              /* :232 */
  }
  iter = iter - 1;                              /* :238 */

done:
  denormalise(A, b, fn);                        /* :255-262 */
  *fin_res = res;
  *niter = iter;
  free(p1); free(q1); free(r1); free(p2); free(q2); free(r2); free(d); free(d_inv);
  return 0;
}

/* ------------------------------------------------------------------------
 * Textbook Krylov methods for non-symmetric systems (not in Native_Mod).
 *
 * The reference's "bicg" is the CG recurrence (or_bicg above).  These are
 * the methods its PETSc path offers by name (KSPBICG, KSPCGS, KSPBCGS,
 * Sources/Process/Cpu/Petsc_Interface.cpp:416-436), restated from Barrett et
 * al., "Templates for the Solution of Linear Systems", SIAM 1994: BiCG
 * (Fig. 2.8), CGS (Fig. 2.9) and BiCGSTAB (Fig. 2.10), each with the
 * preconditioner M applied on the left as Templates writes it.  They share
 * Cg's prologue, stopping test and iteration conventions: normalisation by
 * 1/mean(diag) (Cg.f90:62-84), M from Prec_Form (jacobi or none; the Jacobi
 * M is symmetric, so M^-T = M^-1), the two early exits (Cg.f90:95-123), the
 * stopping measure after every iteration on the recursive residual with the
 * updated x, niter = miter when the loop runs out.  The shadow residual is
 * r~0 = r0.  Breakdown (a zero rho, a zero denominator, omega = 0 or a
 * non-finite measure) returns 1 with niter = the failing iteration.
 *
 * A^T p is formed row by row on the symmetric sparsity pattern of the face
 * matrices: (A^T p)_i = 0 + sum over row i's entries j (CSR order) of
 * a(col_j, i) * p(col_j), the transposed value found in row col_j.
 * ------------------------------------------------------------------------ */

/* tpos[j] = C position of the transposed entry of entry j; -1 if the
 * pattern is not symmetric. */
int or_transpose_positions(const or_matrix* A, int* tpos) {
  for (int i = 0; i < A->nt; i++) {
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
      int k = A->col[j] - 1, found = -1;
      for (int m = A->row[k] - 1; m < A->row[k + 1] - 1; m++)
        if (A->col[m] - 1 == i) { found = m; break; }
      if (found < 0) return -1;
      tpos[j] = found;
    }
  }
  return 0;
}

static void spmv_t(const or_matrix* A, int ni, const int* tpos, const double* p, double* q) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ni; i++) {
    double qi = 0.0;
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
      int k = A->col[j] - 1;
      qi = qi + A->val[tpos[j]] * p[k];
    }
    q[i] = qi;
  }
}

static double dot(const double* a, const double* b, int n) {
  if (reordered()) return blocked(a, b, n, 0);
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}

int or_krylov(or_matrix* A, double* x, double* b, int method, int prec, int miter,
              int* niter, double tol, double* fin_res, const double* norm, int crit,
              double* omegas) {
  const int nt = A->nt, ni = A->ni;
  const size_t nv = 11;
  double* w = (double*)calloc(nv * (size_t)nt, sizeof(double));
  int* tpos = method == OR_BICG_TRUE ? (int*)malloc(sizeof(int) * (size_t)(A->row[nt] - 1)) : NULL;
  if (!w || (method == OR_BICG_TRUE && !tpos)) { free(w); free(tpos); return -2; }
  if (tpos && or_transpose_positions(A, tpos) != 0) { free(w); free(tpos); return -3; }
  double *r = w, *rt = w + nt, *p = w + 2 * (size_t)nt, *q = w + 3 * (size_t)nt,
         *z = w + 4 * (size_t)nt, *u = w + 5 * (size_t)nt, *v = w + 6 * (size_t)nt,
         *t = w + 7 * (size_t)nt, *y = w + 8 * (size_t)nt, *d = w + 9 * (size_t)nt,
         *d_inv = w + 10 * (size_t)nt;
  double res = 0.0, rho = 0.0, rho_old = 1.0, alfa = 0.0, omega = 1.0, beta, bnrm2;
  int iter, status = 0;
  if (prec == OR_PREC_ICC) prec = -1; /* ICC: not offered for these methods */
  if (prec < 0) { free(w); free(tpos); return -4; }

  double fn = normalise(A, b);
  or_prec_form(A, ni, d, d_inv, prec);
  double bnorm = sqrt(sumsq(b, ni));
  if (crit == OR_CRIT_RELRES)
    bnrm2 = bnorm > 0.0 ? 1.0 : 0.0;
  else
    bnrm2 = or_nrms(A, ni, b, x, norm);
  if (bnrm2 < tol) { iter = 0; goto done; }
  or_residual(A, ni, r, b, x);
  res = measure(A, ni, r, x, norm, crit, bnorm);
  if (res < tol) { iter = 0; goto done; }
  for (int i = 0; i < ni; i++) rt[i] = r[i];     /* r~0 = r0 */

  for (iter = 1; iter <= miter; iter++) {
    if (method == OR_BICG_TRUE) {                 /* Templates Fig. 2.8 */
      /* z = M^-1 r (in z), z~ = M^-T r~ (in y) */
      or_prec_solve(A, ni, d, d_inv, z, r, prec);
      or_prec_solve(A, ni, d, d_inv, y, rt, prec);
      rho = dot(z, rt, ni);
      if (rho == 0.0) { status = 1; goto done; }
      if (iter == 1) {
        for (int i = 0; i < ni; i++) { p[i] = z[i]; u[i] = y[i]; }   /* u: p~ */
      } else {
        beta = rho / rho_old;
        for (int i = 0; i < ni; i++) {
          p[i] = z[i] + beta * p[i];
          u[i] = y[i] + beta * u[i];
        }
      }
      or_spmv(A, ni, p, q);                       /* q = A p     */
      spmv_t(A, ni, tpos, u, v);                  /* v = A^T p~  */
      double den = dot(u, q, ni);
      if (den == 0.0) { status = 1; goto done; }
      alfa = rho / den;
      for (int i = 0; i < ni; i++) {
        x[i] = x[i] + alfa * p[i];
        r[i] = r[i] - alfa * q[i];
        rt[i] = rt[i] - alfa * v[i];
      }
    } else if (method == OR_CGS) {                /* Templates Fig. 2.9 */
      rho = dot(rt, r, ni);
      if (rho == 0.0) { status = 1; goto done; }
      if (iter == 1) {
        for (int i = 0; i < ni; i++) { u[i] = r[i]; p[i] = u[i]; }
      } else {
        beta = rho / rho_old;
        for (int i = 0; i < ni; i++) {
          u[i] = r[i] + beta * q[i];
          p[i] = u[i] + beta * (q[i] + beta * p[i]);
        }
      }
      or_prec_solve(A, ni, d, d_inv, z, p, prec); /* p^ (in z) */
      or_spmv(A, ni, z, v);                       /* v^ = A p^ */
      double den = dot(rt, v, ni);
      if (den == 0.0) { status = 1; goto done; }
      alfa = rho / den;
      for (int i = 0; i < ni; i++) {
        q[i] = u[i] - alfa * v[i];
        y[i] = u[i] + q[i];
      }
      or_prec_solve(A, ni, d, d_inv, t, y, prec); /* u^ = M^-1 (u + q) (in t) */
      for (int i = 0; i < ni; i++) x[i] = x[i] + alfa * t[i];
      or_spmv(A, ni, t, y);                       /* q^ = A u^ (in y) */
      for (int i = 0; i < ni; i++) r[i] = r[i] - alfa * y[i];
    } else {                                      /* BiCGSTAB, Fig. 2.10 */
      rho = dot(rt, r, ni);
      if (rho == 0.0) { status = 1; goto done; }
      if (iter == 1) {
        for (int i = 0; i < ni; i++) p[i] = r[i];
      } else {
        beta = (rho / rho_old) * (alfa / omega);
        for (int i = 0; i < ni; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
      }
      or_prec_solve(A, ni, d, d_inv, z, p, prec); /* p^ (in z) */
      or_spmv(A, ni, z, v);                       /* v = A p^ */
      double den = dot(rt, v, ni);
      if (den == 0.0) { status = 1; goto done; }
      alfa = rho / den;
      for (int i = 0; i < ni; i++) {
        u[i] = r[i] - alfa * v[i];                /* s (in u) */
        q[i] = x[i] + alfa * z[i];                /* x + alpha p^ (in q) */
      }
      res = measure(A, ni, u, q, norm, crit, bnorm);
      if (!isfinite(res)) { status = 1; goto done; }
      if (res < tol) {                            /* early exit on s */
        for (int i = 0; i < ni; i++) { x[i] = q[i]; r[i] = u[i]; }
        goto done;
      }
      or_prec_solve(A, ni, d, d_inv, y, u, prec); /* s^ (in y) */
      or_spmv(A, ni, y, t);                       /* t = A s^ */
      double tt = dot(t, t, ni);
      if (tt == 0.0) { status = 1; goto done; }
      omega = dot(t, u, ni) / tt;
      if (omegas) omegas[iter - 1] = omega;
      for (int i = 0; i < ni; i++) {
        x[i] = q[i] + omega * y[i];
        r[i] = u[i] - omega * t[i];
      }
    }
    res = measure(A, ni, r, x, norm, crit, bnorm);
    if (!isfinite(res)) { status = 1; goto done; }
    if (res < tol) goto done;
    if (method == OR_BCGS && omega == 0.0) { status = 1; goto done; }
    rho_old = rho;
  }
  iter = iter - 1;

done:
  denormalise(A, b, fn);
  *fin_res = res;
  *niter = iter;
  free(w);
  free(tpos);
  return status;
}

/* ------------------------------------------------------------------------
 * Pressure-system assembly from face data: A % val = 0 and b = 0
 * (Compute_Pressure.f90:121-122), the matrix lines of the Rhie and Chow face
 * loop over internal faces (Rhie_And_Chow.f90:184-205), the right-hand side
 * from the volume fluxes and the diagonal terms of boundary faces with a
 * pressure / ambient condition (Compute_Pressure.f90:147-169), the boundary
 * regions' b (:171-189: b = 0 in pressure regions, b = fc v_m(c1) pp(c2) in
 * ambient ones, region by region), and Set_Singular's a_ii * (1 + MICRO)
 * (Solver_Mod/Set_Singular.f90:28-31) when `singular`.  Serial, in face order.
 * c1, c2: 1-based cells (c2 <= 0: boundary face); pos: 2 x n_faces 1-based
 * CSR positions of (c1,c2) and (c2,c1); diag_bc: per-face flag; b_region:
 * the type of the face's region (0 none, 1 pressure, 2 ambient); region: the
 * region number of each face (1.., 0 for internal faces): Faces_In_Region
 * are the faces of a number in increasing order; pp: pp % n at cell 1
 * (boundary cell c at pp[c - 1]), read by ambient regions only.
 * ------------------------------------------------------------------------ */
void or_pressure_assemble(int nc, int nnz, int n_faces, const int* c1, const int* c2,
                          const int* pos, const int* dia, const int* diag_bc,
                          const int* b_region, const int* region, const double* fc,
                          const double* fw, const double* vm, const double* vflux,
                          const double* pp, double* val, double* b, int singular) {
  for (int j = 0; j < nnz; j++) val[j] = 0.0;
  for (int i = 0; i < nc; i++) b[i] = 0.0;
  for (int s = 0; s < n_faces; s++) {           /* Rhie_And_Chow.f90:184-205 */
    if (c2[s] <= 0) continue;
    const int a = c1[s] - 1, c = c2[s] - 1;
    const double fs = fw[s];
    const double a12 = fc[s] * (fs * vm[a] + (1.0 - fs) * vm[c]);
    val[pos[2 * s] - 1] = -a12;
    val[pos[2 * s + 1] - 1] = -a12;
    val[dia[a] - 1] = val[dia[a] - 1] + a12;
    val[dia[c] - 1] = val[dia[c] - 1] + a12;
  }
  for (int s = 0; s < n_faces; s++) {           /* Compute_Pressure.f90:147-169 */
    const int a = c1[s] - 1;
    if (c2[s] > 0) {
      b[a] = b[a] - vflux[s];
      b[c2[s] - 1] = b[c2[s] - 1] + vflux[s];
    } else {
      b[a] = b[a] - vflux[s];
      if (diag_bc[s]) {
        const double a12 = fc[s] * vm[a];
        val[dia[a] - 1] = val[dia[a] - 1] + a12;
      }
    }
  }
  int nreg = 0;                                 /* :171-189, Boundary_Regions() */
  for (int s = 0; s < n_faces; s++) nreg = region[s] > nreg ? region[s] : nreg;
  for (int reg = 1; reg <= nreg; reg++)
    for (int s = 0; s < n_faces; s++) {         /* Faces_In_Region(reg) */
      if (region[s] != reg) continue;
      const int a = c1[s] - 1;
      if (b_region[s] == 1) {                   /* :173-178, pressure */
        b[a] = 0.0;
      } else if (b_region[s] == 2) {            /* :180-187, ambient */
        const double a12 = fc[s] * vm[a];
        b[a] = a12 * pp[c2[s] - 1];
      }
    }
  if (singular)                                 /* Set_Singular.f90:28-31 */
    for (int i = 0; i < nc; i++) val[dia[i] - 1] = val[dia[i] - 1] * (1.0 + OR_MICRO);
}

/* Compute_Momentum.f90:201-204: vel_max = max(MICRO, max_c |u_c|) over
 * cells 1..nc (before the reference's Global % Max_Real). */
double or_vel_max(int nc, const double* u, const double* v, const double* w) {
  double vel_max = OR_MICRO;
  for (int c = 0; c < nc; c++) {
    const double m = sqrt(u[c] * u[c] + v[c] * v[c] + w[c] * w[c]);
    vel_max = m > vel_max ? m : vel_max;
  }
  return vel_max;
}

/* ------------------------------------------------------------------------
 * Momentum assembly of component m->component (1, 2, 3 = u, v, w), serial in
 * the reference's loop order.  Cells are 1-based; a boundary cell c < 0 of a
 * [b] array is element c - 1 of the pointer (the pointer is cell 1).
 * ---------------------------------------------------------------------- */
#define OR_FEMTO 1.0e-15
#define CELL(a, c) ((a)[(c) - 1])

static double or_min(double a, double b) { return b < a ? b : a; }
static double or_max(double a, double b) { return b > a ? b : a; }

/* Numerics_Mod/Advection_Scheme.f90:22-183: face value phi_f of face s for
 * every scheme but CENTRAL (which Advection_Term never passes here). */
static double or_adv_scheme(const or_momentum* m, int s, int c1, int c2,
                            const double* pmin, const double* pmax) {
  const double* phi = m->phi_n;
  double fj, sgn;
  int c, d;
  if (m->v_flux[s] > 0.0) {                          /* :43-54 */
    fj = 1.0 - m->f[s]; c = c1; d = c2; sgn = +1.0;
  } else {
    fj = m->f[s]; c = c2; d = c1; sgn = -1.0;
  }
  if (m->adv_scheme == OR_ADV_UPWIND) return CELL(phi, c);              /* :103-105 */
  if (m->adv_scheme == OR_ADV_BLENDED) {                                 /* :107-112 */
    const double phi_f_c = CELL(phi, c) + fj * sgn * (CELL(phi, c2) - CELL(phi, c1));
    const double phi_f_u = CELL(phi, c);
    return m->blend * phi_f_c + (1.0 - m->blend) * phi_f_u;
  }
  const double gd = CELL(m->gx, c) * m->dx[s] + CELL(m->gy, c) * m->dy[s] + CELL(m->gz, c) * m->dz[s];
  double phi_f;
  if (m->adv_scheme == OR_ADV_GAMMA) {                                   /* :117-139 */
    const double beta = 0.1;
    const double phi_star = m->v_flux[s] > 0.0
                                ? 1.0 - (CELL(phi, d) - CELL(phi, c)) / (2.0 * gd)
                                : 1.0 + (CELL(phi, d) - CELL(phi, c)) / (2.0 * gd);
    const double gamma_c = phi_star / beta;
    if (phi_star < beta && phi_star > 0.0)
      phi_f = (1.0 - gamma_c * (1.0 - m->f[s])) * CELL(phi, c) + gamma_c * (1.0 - m->f[s]) * CELL(phi, d);
    else if (phi_star < 1.0 && phi_star >= beta)
      phi_f = m->f[s] * CELL(phi, c) + (1.0 - m->f[s]) * CELL(phi, d);
    else
      phi_f = CELL(phi, c);
    return phi_f;
  }
  const double phi_star = m->v_flux[s] > 0.0 ? CELL(phi, d) - 2.0 * gd    /* :56-64 */
                                             : CELL(phi, d) + 2.0 * gd;
  const double phi_u = or_max(CELL(pmin, c), or_min(phi_star, CELL(pmax, c)));
  double denom = CELL(phi, d) - CELL(phi, c);
  if (fabs(denom) < OR_FEMTO) denom = denom + copysign(OR_FEMTO, denom);
  const double rj = (CELL(phi, c) - phi_u) / denom;
  const double g_d = 0.5 * fj * (1.0 + fj), g_u = 0.5 * fj * (1.0 - fj);
  double phij = fj;                                                      /* :78-101 */
  if (m->adv_scheme == OR_ADV_QUICK) {
    const double alfa = 0.0;
    phij = (g_d - alfa) + (g_u + alfa) * rj;
  } else if (m->adv_scheme == OR_ADV_LUDS) {
    const double alfa = 0.5 * fj * (1 + fj);
    phij = (g_d - alfa) + (g_u + alfa) * rj;
  } else if (m->adv_scheme == OR_ADV_MINMOD) {
    phij = fj * or_max(0.0, or_min(rj, 1.0));
  } else if (m->adv_scheme == OR_ADV_SMART) {
    const double beta1 = 3.0, beta2 = 1.0;
    phij = or_max(0.0, or_min(or_min((beta1 - 1.0) * rj, g_d + g_u * rj), beta2));
  } else if (m->adv_scheme == OR_ADV_AVL_SMART) {
    const double beta1 = 1.0 + fj * (2.0 + fj), beta2 = fj * (2.0 - fj);
    phij = or_max(0.0, or_min(or_min((beta1 - 1.0) * rj, g_d + g_u * rj), beta2));
  } else if (m->adv_scheme == OR_ADV_SUPERBEE) {
    phij = 0.5 * or_max(or_max(0.0, or_min(2.0 * rj, 1.0)), or_min(rj, 2.0));
  }
  return CELL(phi, c) + phij * sgn * (CELL(phi, c2) - CELL(phi, c1));    /* :114 */
}

double or_adv_face(int scheme, double f, double flux, const double* phi, const double* grad,
                   const double* d, const double* pmin, const double* pmax, double blend) {
  const double gx[2] = {grad[0], grad[3]}, gy[2] = {grad[1], grad[4]}, gz[2] = {grad[2], grad[5]};
  or_momentum m;
  memset(&m, 0, sizeof(m));
  m.adv_scheme = scheme;
  m.blend = blend;
  m.f = &f;
  m.v_flux = &flux;
  m.dx = &d[0];
  m.dy = &d[1];
  m.dz = &d[2];
  m.phi_n = phi;
  m.gx = gx;
  m.gy = gy;
  m.gz = gz;
  if (scheme == OR_ADV_CENTRAL) return f * phi[0] + (1.0 - f) * phi[1];
  return or_adv_scheme(&m, 0, 1, 2, pmin, pmax);
}

void or_momentum_assemble(int nc, int nnz, int n_faces, const int* c1, const int* c2,
                          const int* pos, const int* dia, const or_momentum* m, double* val,
                          double* b, double* v_m) {
  const int nb = m->n_bnd;
  double* cross = calloc((size_t)nc, sizeof(double));
  double* fi = calloc((size_t)nc, sizeof(double));
  double* advect = calloc((size_t)nc, sizeof(double));
  double* upwind = calloc((size_t)nc, sizeof(double));
  double* mn = malloc(sizeof(double) * (size_t)(nb + 1 + nc));
  double* mx = malloc(sizeof(double) * (size_t)(nb + 1 + nc));
  double* pmin = mn + nb + 1;       /* cell 1; boundary cells below it */
  double* pmax = mx + nb + 1;
  const double* phi = m->phi_n;
  const double* rho = m->density;
  /* rotated views of component i (Compute_Momentum.f90:58-98) */
  const int i = m->component;
  const double* g[3] = {m->gx, m->gy, m->gz};
  const double* sv[3] = {m->sx, m->sy, m->sz};
  const double* dv[3] = {m->dx, m->dy, m->dz};
  const double *ui_i = g[(i - 1) % 3], *ui_j = g[i % 3], *ui_k = g[(i + 1) % 3];
  const double *si = sv[(i - 1) % 3], *sj = sv[i % 3], *sk = sv[(i + 1) % 3];
  const double *di = dv[(i - 1) % 3], *dj = dv[i % 3], *dk = dv[(i + 1) % 3];

  for (int j = 0; j < nnz; j++) val[j] = 0.0;                 /* :119-123 */
  for (int c = 0; c < nc; c++) b[c] = 0.0;

  /* Advection_Term.f90:35-120 with b = fi */
  if (m->adv_scheme != OR_ADV_CENTRAL) {                      /* Min_Max.f90:19-34 */
    for (int c = -nb; c <= nc; c++)
      if (c != 0) { CELL(pmin, c) = CELL(phi, c); CELL(pmax, c) = CELL(phi, c); }
    for (int s = 0; s < n_faces; s++) {
      const int a = c1[s], e = c2[s];
      CELL(pmin, a) = or_min(CELL(pmin, a), CELL(phi, e));
      CELL(pmin, e) = or_min(CELL(pmin, e), CELL(phi, a));
      CELL(pmax, a) = or_max(CELL(pmax, a), CELL(phi, e));
      CELL(pmax, e) = or_max(CELL(pmax, e), CELL(phi, a));
    }
  }
  for (int pass = 0; pass < 2; pass++) {        /* boundary regions, then domain faces */
    for (int s = 0; s < n_faces; s++) {
      const int a = c1[s], e = c2[s];
      if ((pass == 0) != (e < 0)) continue;
      double phif = m->f[s] * CELL(phi, a) + (1.0 - m->f[s]) * CELL(phi, e);
      if (m->adv_scheme != OR_ADV_CENTRAL) phif = or_adv_scheme(m, s, a, e, pmin, pmax);
      CELL(advect, a) = CELL(advect, a) - m->v_flux[s] * phif * CELL(rho, a);
      if (e > 0) CELL(advect, e) = CELL(advect, e) + m->v_flux[s] * phif * CELL(rho, e);
    }
    if (m->blend_matrix) {
      for (int s = 0; s < n_faces; s++) {
        const int a = c1[s], e = c2[s];
        if ((pass == 0) != (e < 0)) continue;
        const double up = m->v_flux[s] < 0 ? CELL(phi, e) : CELL(phi, a);
        CELL(upwind, a) = CELL(upwind, a) - m->v_flux[s] * up * CELL(rho, a);
        if (e > 0) CELL(upwind, e) = CELL(upwind, e) + m->v_flux[s] * up * CELL(rho, e);
      }
    }
  }
  for (int c = 0; c < nc; c++) fi[c] = fi[c] + advect[c] - upwind[c];   /* :115-117 */

  /* diffusion, Compute_Momentum.f90:238-301 */
  for (int s = 0; s < n_faces; s++) {
    const int a = c1[s], e = c2[s];
    const double w = m->fw[s], vis_eff = m->vis_eff[s];
    const double f_stress = m->f_stress ? m->f_stress[s] : 0.0;
    const double ui_i_f = w * CELL(ui_i, a) + (1.0 - w) * CELL(ui_i, e);
    const double ui_j_f = w * CELL(ui_j, a) + (1.0 - w) * CELL(ui_j, e);
    const double ui_k_f = w * CELL(ui_k, a) + (1.0 - w) * CELL(ui_k, e);
    const double uj_i_f = w * CELL(m->uj_i, a) + (1.0 - w) * CELL(m->uj_i, e);
    const double uk_i_f = w * CELL(m->uk_i, a) + (1.0 - w) * CELL(m->uk_i, e);
    const double ui_si = (ui_i_f + ui_i_f) * si[s] + (ui_j_f + uj_i_f) * sj[s] + (ui_k_f + uk_i_f) * sk[s];
    const double ui_di = ui_i_f * di[s] + ui_j_f * dj[s] + ui_k_f * dk[s];
    const double f_ex = vis_eff * ui_si;
    const double m0 = vis_eff * m->fc[s];
    const double f_im = ui_di * m0;
    CELL(cross, a) = CELL(cross, a) + f_ex - f_im + f_stress * CELL(rho, a);
    if (e > 0) CELL(cross, e) = CELL(cross, e) - f_ex + f_im - f_stress * CELL(rho, e);
    double m12 = m0, m21 = m0;
    if (m->blend_matrix) {
      m12 = m12 - or_min(m->v_flux[s], 0.0) * CELL(rho, a);
      m21 = m21 + or_max(m->v_flux[s], 0.0) * CELL(rho, e);
    }
    if (e > 0) {
      val[pos[2 * s] - 1] = val[pos[2 * s] - 1] - m12;
      val[CELL(dia, a) - 1] = val[CELL(dia, a) - 1] + m12;
      val[pos[2 * s + 1] - 1] = val[pos[2 * s + 1] - 1] - m21;
      val[CELL(dia, e) - 1] = val[CELL(dia, e) - 1] + m21;
    } else if (e < 0 && m->wall[s]) {
      val[CELL(dia, a) - 1] = val[CELL(dia, a) - 1] + m12;
      CELL(fi, a) = CELL(fi, a) + m12 * CELL(phi, e);
    }
  }
  for (int c = 1; c <= nc; c++) CELL(fi, c) = CELL(fi, c) + CELL(cross, c);   /* :306-308 */

  for (int c = 1; c <= nc; c++) {                             /* Inertial_Term.f90:27-41 */
    const double a0 = CELL(rho, c) * CELL(m->vol, c) / m->dt;
    double* dd = &val[CELL(dia, c) - 1];
    if (m->td_scheme == OR_TD_LINEAR) {
      *dd = *dd + a0;
      CELL(fi, c) = CELL(fi, c) + a0 * CELL(m->phi_o, c);
    } else if (m->td_scheme == OR_TD_PARABOLIC) {
      *dd = *dd + 1.5 * a0;
      CELL(fi, c) = CELL(fi, c) + 2.0 * a0 * CELL(m->phi_o, c) - 0.5 * a0 * CELL(m->phi_oo, c);
    }
  }
  for (int c = 1; c <= nc; c++) {                  /* Compute_Momentum.f90:322-350 */
    CELL(fi, c) = CELL(fi, c) + m->p_drop * CELL(m->vol, c);
    const double cf = m->cell_f ? CELL(m->cell_f, c) : 0.0;
    CELL(fi, c) = CELL(fi, c) + cf * CELL(m->vol, c);
    CELL(b, c) = CELL(fi, c) - CELL(m->p_i, c) * CELL(m->vol, c);
    if (m->st) CELL(b, c) = CELL(b, c) + CELL(m->st, c) * CELL(m->vol, c);
  }
  for (int c = 1; c <= nc; c++)                                 /* :361-363 */
    CELL(v_m, c) = CELL(m->vol, c) / val[CELL(dia, c) - 1];
  for (int c = 1; c <= nc; c++) {                               /* Under_Relax.f90:18-24 */
    double* dd = &val[CELL(dia, c) - 1];
    *dd = *dd / m->urf;
    CELL(b, c) = CELL(b, c) + *dd * (1.0 - m->urf) * CELL(phi, c);
  }
  free(cross); free(fi); free(advect); free(upwind); free(mn); free(mx);
}

/* ------------------------------------------------------------------------
 * Rhie-Chow volume fluxes, Rhie_And_Chow.f90:86-259 (the matrix lines of
 * the same face loop are or_pressure_assemble's).  Cells 1-based.
 * ---------------------------------------------------------------------- */
void or_rhie_chow_fluxes(int nc, int n_faces, const int* c1, const int* c2,
                         const or_rhie_chow* r, double* v_flux) {
  const int vof = r->vof && r->sigma > 1.0e-15;      /* TINY = FEMTO, :104-105 */
  double* pst = malloc(sizeof(double) * 3 * (size_t)nc);
  double* uc = malloc(sizeof(double) * 3 * (size_t)nc);
  double* tm = malloc(sizeof(double) * (size_t)nc);
  for (int c = 0; c < nc; c++) {                              /* :86-91, 107-112 */
    pst[3 * c] = r->px[c];
    pst[3 * c + 1] = r->py[c];
    pst[3 * c + 2] = r->pz[c];
    if (vof) {
      pst[3 * c] = pst[3 * c] - r->sigma * r->curv[c] * r->col_x[c];
      pst[3 * c + 1] = pst[3 * c + 1] - r->sigma * r->curv[c] * r->col_y[c];
      pst[3 * c + 2] = pst[3 * c + 2] - r->sigma * r->curv[c] * r->col_z[c];
    }
    uc[3 * c] = r->u[c];                                      /* :130-134 */
    uc[3 * c + 1] = r->v[c];
    uc[3 * c + 2] = r->w[c];
    if (r->choi) {                                            /* :139-163 */
      tm[c] = r->v_m[c] * r->density[c] / r->dt;
      uc[3 * c] = uc[3 * c] - (r->w_o * r->u_o[c] + r->w_oo * r->u_oo[c]) * tm[c];
      uc[3 * c + 1] = uc[3 * c + 1] - (r->w_o * r->v_o[c] + r->w_oo * r->v_oo[c]) * tm[c];
      uc[3 * c + 2] = uc[3 * c + 2] - (r->w_o * r->w_o_[c] + r->w_oo * r->w_oo_[c]) * tm[c];
    }
    if (r->gu) {                                              /* :169-179 */
      uc[3 * c] = uc[3 * c] - r->v_m[c] * r->cell_fx[c];
      uc[3 * c + 1] = uc[3 * c + 1] - r->v_m[c] * r->cell_fy[c];
      uc[3 * c + 2] = uc[3 * c + 2] - r->v_m[c] * r->cell_fz[c];
    }
  }
  for (int s = 0; s < n_faces; s++) {                         /* :184-259 */
    if (c2[s] <= 0) continue;
    const int a = c1[s] - 1, b = c2[s] - 1;
    const double fs = r->f[s];
    double pst_d = r->p[a] - r->p[b];                         /* :94-98 */
    if (vof) {                                                /* :115-123 */
      const double curv_f = fs * r->curv[a] + (1.0 - fs) * r->curv[b];
      pst_d = pst_d - r->sigma * curv_f * (r->col[a] - r->col[b]);
    }
    const double u_f = fs * uc[3 * a] + (1.0 - fs) * uc[3 * b];
    const double v_f = fs * uc[3 * a + 1] + (1.0 - fs) * uc[3 * b + 1];
    const double w_f = fs * uc[3 * a + 2] + (1.0 - fs) * uc[3 * b + 2];
    const double a12 = r->fc[s] * (fs * r->v_m[a] + (1.0 - fs) * r->v_m[b]);
    const double px_f = (fs * (pst[3 * a] * r->v_m[a]) + (1.0 - fs) * (pst[3 * b] * r->v_m[b])) * r->dx[s];
    const double py_f = (fs * (pst[3 * a + 1] * r->v_m[a]) + (1.0 - fs) * (pst[3 * b + 1] * r->v_m[b])) * r->dy[s];
    const double pz_f = (fs * (pst[3 * a + 2] * r->v_m[a]) + (1.0 - fs) * (pst[3 * b + 2] * r->v_m[b])) * r->dz[s];
    double fl = (u_f * r->sx[s] + v_f * r->sy[s] + w_f * r->sz[s]) + a12 * pst_d +
                r->fc[s] * (px_f + py_f + pz_f);
    if (r->choi)                                              /* :235-246 */
      fl = fl + (fs * tm[a] + (1.0 - fs) * tm[b]) * (r->w_o * r->flux_o[s] + r->w_oo * r->flux_oo[s]);
    if (r->gu)                                                /* :250-256 */
      fl = fl + (fs * r->v_m[a] + (1.0 - fs) * r->v_m[b]) *
                    (r->face_fx[s] * r->sx[s] + r->face_fy[s] * r->sy[s] + r->face_fz[s] * r->sz[s]);
    v_flux[s] = fl;
  }
  free(pst); free(uc); free(tm);
}

/* ==========================================================================
 * FP32 sweeps inside FP64 refinement: the library's TFG_FP32_IR form.  No
 * counterpart in the reference, which solves in FP64 only; restated here so
 * that the GPU's FP32 product is checked bit for bit and its solves step by
 * step.  The FP64 parts (normalisation, prologue, true residual, stopping
 * test) are Cg.f90's, as or_cg_ex above.
 * ========================================================================== */

/* Neumaier's compensated sum: the running sum and the rounding errors of
 * its additions, added at the end. */
typedef struct {
  double s, c;
} or_csum;

static void csum_add(or_csum* a, double v) {
  const double t = a->s + v;
  if (fabs(a->s) >= fabs(v)) a->c = a->c + ((a->s - t) + v);
  else a->c = a->c + ((v - t) + a->s);
  a->s = t;
}

/* Row sum s_i = sum_j a_ij fn of row i (the products rounded in FP64),
 * compensated, in CSR order: the row's exact sum to within an FP64
 * rounding, even where it cancels to the rounding level of the entries. */
static double row_excess(const or_matrix* A, int i, double fn) {
  or_csum a = {0.0, 0.0};
  for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) csum_add(&a, A->val[j] * fn);
  return a.s + a.c;
}

/* FP32 matrix of the row-excess form, from the FP64 values scaled by fn:
 * the diagonal slot of row i holds fl32(s_i), every other slot
 * fl32(a_ij fn).  Rows 0..ni-1. */
void or_lap32_form(const or_matrix* A, int ni, double fn, float* v32) {
  for (int i = 0; i < ni; i++) {
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) v32[j] = (float)(A->val[j] * fn);
    v32[A->dia[i] - 1] = (float)row_excess(A, i, fn);
  }
}

/* y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i), FP32, terms in CSR order
 * (diagonal first).  Exact in exact arithmetic (it is A x rewritten); in
 * FP32 a vector constant over a row's cells costs no rounding beyond s_i's,
 * so the nearly singular pressure systems keep their small eigenvalues. */
void or_lap32_spmv(const or_matrix* A, int ni, const float* v32, const float* x, float* y) {
  for (int i = 0; i < ni; i++) {
    const float xi = x[i];
    float acc = v32[A->dia[i] - 1] * xi;
    for (int j = A->row[i] - 1; j < A->row[i + 1] - 1; j++) {
      const int k = A->col[j] - 1;
      if (k == i) continue;
      acc = acc + v32[j] * (x[k] - xi);
    }
    y[i] = acc;
  }
}

static double dot32(const float* a, const float* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + (double)a[i] * (double)b[i];
  return s;
}

/* z = fl32(M^-1 fl64(r)) of the ICC (DIC) of the scaled matrix, the FP32
 * form's ICC: the FP64 sweeps of Prec_Solve.f90:64-92 on the widened FP32
 * residual, rounded back. */
static void ir_icc_z(const or_matrix* A, int ni, const double* d, const double* dinv,
                     const float* r, double* rw, double* zw, float* z) {
  for (int i = 0; i < ni; i++) rw[i] = (double)r[i];
  or_prec_solve(A, ni, d, dinv, zw, rw, OR_PREC_ICC);
  for (int i = 0; i < ni; i++) z[i] = (float)zw[i];
}

/* FP32 Jacobi / ICC / unpreconditioned CG in the row-excess form with FP64
 * reliable updates (ICC: z from ir_icc_z); with `deflate`, the constant vector w is deflated
 * (deflated PCG, Saad, Yeung, Erhel and Guyomarc'h 2000): E = w'Aw =
 * sum_i s_i, every direction p = z + beta p - w (s'z) / E, and every
 * restart begins with the coarse correction x += c w, c = w'(b - A x) / E
 * computed as (sum_i b_i - sum_i s_i x_i) / E: with the compensated row
 * sums s_i and a compensated sum of b this is accurate, where summing the
 * residual itself would carry its rows' rounding errors, which 1 / E
 * amplifies on a nearly singular matrix.  The Krylov vectors r, p, q, the
 * correction e and the product are FP32; alpha, beta and every dot product
 * are FP64 (each FP32 product is exact in FP64).  Every `check` iterations
 * (and at miter) the FP32 residual is looked at; a refinement follows when
 * it has dropped by `delta` since the last restart, when its estimate of
 * the stopping measure is below tol, when `force` iterations (> 0) have
 * passed, or at miter: x += e in FP64, the FP64 true residual b - A x, the
 * stopping test on it, and the FP32 CG restarts from it.  niter counts FP32
 * iterations; fin_res is the measure of the last FP64 true residual.
 * Returns 0, 1 on breakdown ((p, q) = 0 or not finite: x keeps the
 * correction so far), -2 out of memory. */
int or_ir_cg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
             double tol, double* fin_res, const double* norm, int crit, double delta,
             int check, int force, int deflate, int* refinements) {
  const int nt = A->nt, ni = A->ni;
  const long long nnz = A->row[nt] - 1;
  double* r64 = (double*)calloc((size_t)nt, sizeof(double));
  double* s64 = (double*)calloc((size_t)nt, sizeof(double));
  float* v32 = (float*)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
  float* rd = (float*)malloc(sizeof(float) * (size_t)nt);
  float* r = (float*)calloc((size_t)nt, sizeof(float));
  float* z = (float*)calloc((size_t)nt, sizeof(float));
  float* p = (float*)calloc((size_t)nt, sizeof(float));
  float* q = (float*)calloc((size_t)nt, sizeof(float));
  float* e = (float*)calloc((size_t)nt, sizeof(float));
  const int icc = prec == OR_PREC_ICC;
  double* dic = (double*)calloc((size_t)(icc ? 4 * nt : 1), sizeof(double));
  if (!r64 || !s64 || !v32 || !rd || !r || !z || !p || !q || !e || !dic) {
    free(r64); free(s64); free(v32); free(rd); free(r); free(z); free(p); free(q); free(e);
    free(dic);
    return -2;
  }
  double *dic_d = dic, *dic_dinv = dic + nt, *rw = dic + 2 * nt, *zw = dic + 3 * nt;
  double res = 0.0, rho = 0.0, rref = 0.0;
  int iter = 0, nref = 0, brk = 0;
  if (check < 1) check = 1;

  const double fn = normalise(A, b);
  if (icc) or_prec_form(A, ni, dic_d, dic_dinv, OR_PREC_ICC);   /* Cg.f90:89 */
  or_lap32_form(A, ni, 1.0, v32);             /* values already scaled */
  double E = 0.0;
  or_csum bsum = {0.0, 0.0};
  for (int i = 0; i < ni; i++) {
    rd[i] = (float)(1.0 / A->val[A->dia[i] - 1]);
    s64[i] = row_excess(A, i, 1.0);
    E = E + s64[i];
    csum_add(&bsum, b[i]);
  }
  if (!(E > 0.0)) deflate = 0;
  const double bnorm = sqrt(sumsq(b, ni));
  const double bnrm2 = crit == OR_CRIT_RELRES ? (bnorm > 0.0 ? 1.0 : 0.0)
                                              : or_nrms(A, ni, b, x, norm);
  if (bnrm2 < tol) goto done;
  or_residual(A, ni, r64, b, x);
  res = measure(A, ni, r64, x, norm, crit, bnorm);
  if (res < tol) goto done;

  {
    int klast = 0;
    double xmin = 0.0, xmax = 0.0;
    for (int k = 0; k <= miter; k++) {
      if (k == 0 || klast == k) {             /* start / restart from r64 */
        xmin = HUGE_VAL;                      /* x range of the true residual */
        xmax = -HUGE_VAL;
        double sx = 0.0;
        for (int i = 0; i < ni; i++) {
          xmin = fmin(xmin, x[i]);
          xmax = fmax(xmax, x[i]);
          sx = sx + s64[i] * x[i];
        }
        double c = 0.0;
        if (deflate) {                        /* x += c w: then w'r = 0 */
          c = ((bsum.s - sx) + bsum.c) / E;
          for (int i = 0; i < ni; i++) x[i] = x[i] + c;
        }
        for (int i = 0; i < ni; i++) {        /* r - c A w = r - c s */
          r[i] = (float)(r64[i] - c * s64[i]);
          z[i] = prec == OR_PREC_JACOBI ? r[i] * rd[i] : r[i];
        }
        if (icc) ir_icc_z(A, ni, dic_d, dic_dinv, r, rw, zw, z);
        float mu = 0.0f;                      /* p = z - w (s'z) / E */
        if (deflate) {
          double g = 0.0;
          for (int i = 0; i < ni; i++) g = g + (double)(float)s64[i] * (double)z[i];
          mu = (float)(g / E);
        }
        for (int i = 0; i < ni; i++) p[i] = z[i] - mu;
        rref = dot32(r, r, ni);
        rho = dot32(z, r, ni);
        continue;
      }
      or_lap32_spmv(A, ni, v32, p, q);
      const double pq = dot32(p, q, ni);
      if (pq == 0.0 || !isfinite(pq)) {       /* breakdown: keep the correction */
        for (int i = 0; i < ni; i++) x[i] = x[i] + (double)e[i];
        iter = k;
        brk = 1;
        goto done;
      }
      const double alpha = rho / pq;
      const float a32 = (float)alpha;
      for (int i = 0; i < ni; i++) {
        e[i] = e[i] + a32 * p[i];
        r[i] = r[i] - a32 * q[i];
        z[i] = prec == OR_PREC_JACOBI ? r[i] * rd[i] : r[i];
      }
      if (icc) ir_icc_z(A, ni, dic_d, dic_dinv, r, rw, zw, z);
      const double rz = dot32(z, r, ni), rr = dot32(r, r, ni);
      iter = k;
      if (k % check == 0 || k == miter) {
        double est;
        if (crit == OR_CRIT_RELRES) {
          est = sqrt(rr) / bnorm;
        } else {
          double s = 0.0;
          for (int i = 0; i < ni; i++) {
            const double t = (double)(r[i] * rd[i]);
            s = s + t * t;
          }
          const double range = norm ? fmax(*norm - 0.0, OR_TINY) : fmax(xmax - xmin, OR_TINY);
          est = sqrt(s) / range;
        }
        if (rr <= delta * delta * rref || est < tol || (force > 0 && k - klast >= force) ||
            k == miter || !isfinite(rr)) {
          for (int i = 0; i < ni; i++) {
            x[i] = x[i] + (double)e[i];
            e[i] = 0.0f;
          }
          or_residual(A, ni, r64, b, x);
          res = measure(A, ni, r64, x, norm, crit, bnorm);
          nref++;
          if (res < tol || !isfinite(res) || k == miter) goto done;
          klast = k;
          k--;                                /* restart at the top, same k */
          continue;
        }
      }
      const float beta = (float)(rz / rho);
      rho = rz;
      float mu = 0.0f;
      if (deflate) {
        double g = 0.0;
        for (int i = 0; i < ni; i++) g = g + (double)(float)s64[i] * (double)z[i];
        mu = (float)(g / E);
      }
      for (int i = 0; i < ni; i++) p[i] = (z[i] + beta * p[i]) - mu;
    }
  }
done:
  denormalise(A, b, fn);
  *fin_res = res;
  *niter = iter;
  if (refinements) *refinements = nref;
  free(r64); free(s64); free(v32); free(rd); free(r); free(z); free(p); free(q); free(e);
  free(dic);
  return brk ? 1 : 0;
}
