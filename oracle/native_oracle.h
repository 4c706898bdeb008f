/*
 * CPU restatement of T-Flows' native Krylov solvers (Native_Mod).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against and the CPU baseline that is timed beside it.  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg may load it; the
 * GPU library never calls it and has no CPU fallback.
 *
 * Arrays use C indexing: element i here is Fortran element i+1 of the
 * reference.  Index arrays (row, col, dia) keep the reference's 1-based
 * VALUES, exactly as Matrix_Type stores them.
 */
#ifndef TFLOWS_NATIVE_ORACLE_H
#define TFLOWS_NATIVE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_PREC_NONE = 0, OR_PREC_JACOBI = 1, OR_PREC_ICC = 2 };
enum { OR_CRIT_NRMS = 0, OR_CRIT_RELRES = 1 };

typedef struct {
  int ni;          /* rows solved (cells minus buffer cells)          */
  int nt;          /* rows stored (all cells)                          */
  const int* row;  /* nt+1, 1-based positions                          */
  const int* col;  /* nnz,  1-based cell indices                       */
  const int* dia;  /* nt,   1-based position of the diagonal           */
  double* val;     /* nnz,  scaled and de-scaled in place by the solve */
} or_matrix;

double or_nrms(const or_matrix* A, int ni, const double* r, const double* x,
               const double* norm);
void or_residual(const or_matrix* A, int ni, double* r, const double* b,
                 const double* x);
void or_spmv(const or_matrix* A, int ni, const double* p, double* q);
void or_prec_form(const or_matrix* A, int ni, double* d, double* d_inv, int prec);
void or_prec_solve(const or_matrix* A, int ni, const double* d,
                   const double* d_inv, double* x, const double* b, int prec);

/* Cg.f90 / Bicg.f90.  x: nt entries (x0 in, solution out); b: nt entries.
 * norm: NULL when the reference's optional argument is absent. */
int or_cg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
          double tol, double* fin_res, const double* norm);
int or_bicg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
            double tol, double* fin_res, const double* norm);

/* Same CG recurrence with a selectable stopping test: OR_CRIT_NRMS is the
 * reference's; OR_CRIT_RELRES stops on ||r||_2 / ||b||_2 < tol (recursive
 * residual, in the scaled system, which has the same ratio). */
int or_cg_ex(or_matrix* A, double* x, double* b, int prec, int miter,
             int* niter, double tol, double* fin_res, const double* norm,
             int criterion);

/* Textbook BiCG / CGS / BiCGSTAB (Templates, Barrett et al. 1994) with Cg's
 * prologue, stopping test and conventions; prec: jacobi or none.  Returns 0,
 * 1 on breakdown (niter = failing iteration), -2 out of memory, -3 pattern
 * not symmetric (BiCG needs A^T on the same pattern), -4 unsupported prec.
 * omegas (may be NULL) receives BiCGSTAB's omega of every iteration. */
enum { OR_BICG_TRUE = 1, OR_CGS = 2, OR_BCGS = 3 };
int or_krylov(or_matrix* A, double* x, double* b, int method, int prec, int miter,
              int* niter, double tol, double* fin_res, const double* norm, int crit,
              double* omegas);
int or_transpose_positions(const or_matrix* A, int* tpos);

/* Pressure-system assembly from face data (Rhie_And_Chow.f90:184-205,
 * Compute_Pressure.f90:121-178, Set_Singular.f90:28-31), serial in face
 * order; see native_oracle.c. */
void or_pressure_assemble(int nc, int nnz, int n_faces, const int* c1, const int* c2,
                          const int* pos, const int* dia, const int* diag_bc,
                          const int* b_region, const int* region, const double* fc,
                          const double* fw, const double* vm, const double* vflux,
                          const double* pp, double* val, double* b, int singular);
double or_vel_max(int nc, const double* u, const double* v, const double* w);

/* Momentum assembly of one velocity component from face data
 * (Compute_Momentum.f90:119-366 with Numerics_Mod Advection_Term, Min_Max,
 * Advection_Scheme, Inertial_Term, Under_Relax), serial in the reference's
 * loop order; see native_oracle.c.  Cell arrays point at cell 1; those
 * marked [b] also hold the boundary cells c = -n_bnd..-1 at offset c - 1. */
enum { OR_ADV_CENTRAL = 0, OR_ADV_UPWIND = 1, OR_ADV_LUDS = 2, OR_ADV_QUICK = 3,
       OR_ADV_SMART = 4, OR_ADV_GAMMA = 5, OR_ADV_MINMOD = 6, OR_ADV_BLENDED = 7,
       OR_ADV_SUPERBEE = 8, OR_ADV_AVL_SMART = 9 };
enum { OR_TD_LINEAR = 0, OR_TD_PARABOLIC = 1, OR_TD_NONE = 2 };
typedef struct {
  int component, n_bnd, adv_scheme, blend_matrix, td_scheme;
  double blend, dt, urf, p_drop;
  const double *f, *fw, *fc, *v_flux, *vis_eff, *f_stress;   /* per face; f_stress may be NULL */
  const double *sx, *sy, *sz, *dx, *dy, *dz;                  /* per face */
  const int* wall;                                            /* per face */
  const double *phi_n, *phi_o, *phi_oo, *gx, *gy, *gz, *uj_i, *uk_i, *density, *vol, *p_i,
      *cell_f, *st;  /* [b]: phi_n gx gy gz uj_i uk_i density; cell_f, st may be NULL */
} or_momentum;
void or_momentum_assemble(int nc, int nnz, int n_faces, const int* c1, const int* c2,
                          const int* pos, const int* dia, const or_momentum* m, double* val,
                          double* b, double* v_m);
/* Face value phi_f of one face between cells 1 (c1) and 2 (c2) by the
 * advection-scheme routine of the momentum assembly (the limiter alone, for
 * its tests): f = Grid % f(s), flux = v_flux(s), phi[2] the cell values,
 * grad[6] the two cells' gradients (x, y, z of cell 1, then of cell 2),
 * d[3] = Grid % d{x,y,z}(s), pmin[2] / pmax[2] the cells' neighbour bounds,
 * blend = phi % blends(1) (BLENDED only). */
double or_adv_face(int scheme, double f, double flux, const double* phi, const double* grad,
                   const double* d, const double* pmin, const double* pmax, double blend);

/* Rhie-Chow volume fluxes of the domain faces (Rhie_And_Chow.f90:77-259),
 * serial; cell arrays 1-based at cell 1.  Same fields as tfg_rhie_chow. */
typedef struct {
  const double *f, *fc, *sx, *sy, *sz, *dx, *dy, *dz;
  const double *u, *v, *w, *p, *px, *py, *pz, *v_m;
  int choi;
  const double *u_o, *u_oo, *v_o, *v_oo, *w_o_, *w_oo_, *density, *flux_o, *flux_oo;
  double dt, w_o, w_oo;
  int gu;
  const double *cell_fx, *cell_fy, *cell_fz, *face_fx, *face_fy, *face_fz;
  int vof;
  double sigma;
  const double *curv, *col, *col_x, *col_y, *col_z;
} or_rhie_chow;
void or_rhie_chow_fluxes(int nc, int n_faces, const int* c1, const int* c2,
                         const or_rhie_chow* r, double* v_flux);

/* 0: the reference's serial summation (default); 1, 2: blocked-pairwise
 * summation (blocks of 64, 1000) of every global reduction, to measure
 * rounding sensitivity. */
void or_set_sum_order(int order);
void or_set_progress(int every);
void or_set_parts(const int* part);

/* FP32 sweeps inside FP64 refinement (the library's TFG_FP32_IR form; no
 * counterpart in the reference).  or_lap32_form: the FP32 row-excess matrix
 * from val * fn; or_lap32_spmv: its product; or_ir_cg: the solve (see the
 * definition). */
void or_lap32_form(const or_matrix* A, int ni, double fn, float* v32);
void or_lap32_spmv(const or_matrix* A, int ni, const float* v32, const float* x, float* y);
int or_ir_cg(or_matrix* A, double* x, double* b, int prec, int miter, int* niter,
             double tol, double* fin_res, const double* norm, int crit, double delta,
             int check, int force, int deflate, int* refinements);

#ifdef __cplusplus
}
#endif
#endif
