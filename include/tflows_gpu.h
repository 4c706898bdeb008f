/*
 * tflows_gpu: MI355X (gfx950) Krylov solvers for T-Flows' CSR Matrix_Type.
 *
 * C ABI of the drop-in replacement for the native solver call
 *   Nat % Solve_Native(solver, prec, A, x, b, miter, niter, tol, fin_res, norm)
 * (Sources/Process/Cpu/Native_Mod/Solve_Native.f90:2-46), reached from
 * Solver_Mod % Run (Sources/Process/Cpu/Solver_Mod/Run.f90:23-33).
 *
 * Plain pointers and sizes only.  Index arrays are the reference's 1-based
 * Matrix_Type arrays: row(nt+1), col(nnz), dia(nt)
 * (Sources/Process/Cpu/Matrix_Mod.f90:57-76); the diagonal must be the first
 * entry of every row, as Create_Matrix.f90:112-119 guarantees.
 *
 * Every call returns TFG_OK (0) or an error code; tfg_last_error() gives the
 * message of the last failure on the calling thread.
 */
#ifndef TFLOWS_GPU_H
#define TFLOWS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TFG_OK 0
#define TFG_EINVAL 1      /* invalid argument: unknown solver, bad sizes, ... */
#define TFG_EHIP 2        /* HIP runtime failure                              */
#define TFG_ENOMEM 3      /* device or host allocation failed                 */
#define TFG_EBREAKDOWN 4  /* non-finite residual or zero curvature (p.Ap = 0) */
#define TFG_ENOTSUP 5     /* combination not supported (e.g. MPI buffer cells)*/

/* ---- options of the extended solve ------------------------------------- */
#define TFG_FP64 0        /* all arithmetic in FP64 (the reference's)          */
#define TFG_FP32_IR 1     /* FP32 inner CG inside FP64 iterative refinement   */

#define TFG_CRIT_NRMS 0   /* Normalized_Root_Mean_Square < tol (reference)     */
#define TFG_CRIT_RELRES 1 /* ||b - A x||_2 / ||b||_2 < tol                     */

#define TFG_DEFLATE_AUTO 0  /* FP32_IR: deflate the constant vector when the   */
                            /* values are symmetric and nearly singular       */
#define TFG_DEFLATE_OFF 1
#define TFG_DEFLATE_ON 2

#define TFG_PATH_AUTO 0        /* graph below 2M nonzeros, kernels above       */
#define TFG_PATH_KERNELS 1     /* one launch per phase, device-side stop flag  */
#define TFG_PATH_GRAPH 2       /* chunks of iterations replayed as a hipGraph  */
                               /* (FP64 cg/bicg, jacobi/none; batches too)     */

typedef struct {
  int precision;      /* TFG_FP64 | TFG_FP32_IR                                */
  int criterion;      /* TFG_CRIT_NRMS | TFG_CRIT_RELRES                       */
  int path;           /* TFG_PATH_*                                            */
  int inner_max;      /* FP32_IR: > 0: a refinement at least every inner_max  */
                      /* FP32 iterations (0: only when due, see inner_tol)     */
  double inner_tol;   /* FP32_IR: refinement once the FP32 residual fell by   */
                      /* this factor since the last one (0: 1e-3)              */
  int timing;         /* 1: time sampled SpMV launches with hipEvents          */
  int true_residual;  /* 1: recompute ||b - A x||_2/||b||_2 in FP64 at exit    */
  int textbook_bicg;  /* 1: "bicg" is textbook BiCG (with A^T); 0: the        */
                      /* reference's Bicg.f90, which equals CG                 */
  int deflation;      /* FP32_IR: TFG_DEFLATE_AUTO | _OFF | _ON               */
} tfg_options;

typedef struct {
  int status;           /* same as the return code                            */
  int iterations;       /* Krylov iterations (FP32_IR: inner iterations total)*/
  int refinements;      /* FP32_IR: outer refinement steps                    */
  double true_relres;   /* when requested, else -1                            */
  double solve_ms;      /* device time of the solve (hipEvents)               */
  double spmv_ms;       /* mean duration of the sampled SpMV launches         */
  int spmv_samples;     /* number of SpMV launches timed                      */
  int deflated;         /* FP32_IR: 1 when the constant vector was deflated   */
} tfg_report;

typedef struct tfg_matrix tfg_matrix;

/* Upload the CSR structure once per grid.  Replaces the device side of
 * Native_Type % Create_Native / Matrix_Type % Create_Matrix
 * (Native_Mod/Create_Native.f90:14-21, Matrix_Mod/Create_Matrix.f90:40-195).
 * ni: rows solved (n_cells - n_buff_cells); nt: rows stored (n_cells);
 * nnz = row[nt] - 1.  ni < nt (an MPI rank's matrix) needs the hooks of
 * tfg_matrix_set_comm before a solve. */
int tfg_matrix_create(int ni, int nt, int nnz, const int* row, const int* col,
                      const int* dia, tfg_matrix** handle);
int tfg_matrix_destroy(tfg_matrix* handle);

/* Drop-in for Solve_Native (Native_Mod/Solve_Native.f90:2-46), host arrays.
 * solver: "cg" | "bicg" (Fortran blank padding allowed; len < 0: strlen);
 * beyond the reference, the textbook methods of its PETSc path
 * (Petsc_Interface.cpp:416-436): "cgs", "bcgs" (alias "bicgstab"), and
 * textbook "bicg" via tfg_options.textbook_bicg -- jacobi or none, in FP64 or
 * the FP32 form, with the same normalisation, stopping test and niter/fin_res
 * conventions;
 * prec: "jacobi" | "icc" | anything else = none (Prec_Form.f90:47-89);
 * "icc" (the reference's default) with cg / bicg, in FP64 or in the FP32
 * form (TFG_FP32_IR: z = M^-1 r from the FP64 DIC sweeps of the FP32
 * residual, rounded to FP32).
 * val(nnz), x(nt) = x0 on entry / solution on exit, b(nt).  val and b are
 * left unchanged (the reference scales and de-scales them in place,
 * Cg.f90:77-84, 245-250).  norm: NULL when the optional argument is absent.
 * niter, fin_res: as the reference returns them (Cg.f90:226-253). */
int tfg_solve_native(tfg_matrix* h, const char* solver, int solver_len,
                     const char* prec, int prec_len, const double* val,
                     double* x, double* b, int miter, int* niter, double tol,
                     double* fin_res, const double* norm);

/* Same call with device-resident arrays (device pointers) and options.
 * device_ptrs = 1: val, x, b are device pointers on the handle's device;
 * device_ptrs = 0: host pointers as in tfg_solve_native.  opt/rep may be
 * NULL (reference semantics, no report). */
int tfg_solve_native_ex(tfg_matrix* h, const char* solver, int solver_len,
                        const char* prec, int prec_len, const double* val,
                        double* x, double* b, int miter, int* niter, double tol,
                        double* fin_res, const double* norm, int device_ptrs,
                        const tfg_options* opt, tfg_report* rep);

/* Building blocks on device pointers (the product loop Cg.f90:177-185 and
 * Residual_Vector.f90:34-42, without scaling).  Stream-ordered on the
 * handle's stream; call tfg_synchronize() before reading results. */
int tfg_spmv(tfg_matrix* h, const double* d_val, const double* d_p, double* d_q);
int tfg_residual(tfg_matrix* h, const double* d_val, const double* d_b,
                 const double* d_x, double* d_r);

/* Building block of the FP32 refinement form (no counterpart in the
 * reference): the FP32 product in the row-excess form,
 *   y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i),
 * of the FP32 matrix made from d_val scaled by fn: s_i = fl32(sum_j a_ij fn)
 * (each product rounded in FP64, summed in CSR order in FP64) and
 * fl32(a_ij fn) off the diagonal; the terms are added in CSR order.  d_x,
 * d_y: FP32 device vectors of nt entries.  Runs the symmetric FP32 engine
 * when the matrix offers the symmetric engine and d_val is symmetric bit for
 * bit (then d_x and d_y must be 16-byte aligned), else a CSR engine; not on
 * an internally ordered matrix or with buffer cells (TFG_ENOTSUP).
 * Stream-ordered. */
int tfg_spmv_fp32(tfg_matrix* h, const double* d_val, double fn, const float* d_x, float* d_y);

/* z = M^-1 r with the reference's preconditioner of the unscaled matrix
 * (Prec_Form.f90 + Prec_Solve.f90): "icc" (DIC, level-scheduled, bit-exact
 * to the serial sweeps), "jacobi", anything else = copy.  Stream-ordered. */
int tfg_prec_apply(tfg_matrix* h, const char* prec, int prec_len,
                   const double* d_val, const double* d_r, double* d_z);

/* Time `reps` back-to-back SpMV launches with hipEvents on the handle's
 * stream; *ms_avg = mean duration of one launch. */
int tfg_spmv_bench(tfg_matrix* h, const double* d_val, const double* d_p,
                   double* d_q, int reps, double* ms_avg);

/* Device memory helpers (bytes; copies are synchronous on the handle-free
 * default path). */
int tfg_device_alloc(void** dptr, long long bytes);
int tfg_device_free(void* dptr);
int tfg_memcpy_h2d(void* dst, const void* src, long long bytes);
int tfg_memcpy_d2h(void* dst, const void* src, long long bytes);
int tfg_synchronize(tfg_matrix* h);

/* Measurement helper with no counterpart in the reference: a plain 16-byte
 * streaming copy of `bytes` (multiple of 16, 16-byte aligned device
 * pointers), launched reps+1 times on the default stream; *ms_avg = mean of
 * the last reps.  Its known traffic calibrates the HBM byte counters. */
int tfg_stream_copy(void* d_dst, const void* d_src, long long bytes, int reps, double* ms_avg);
/* The same with mode 0 copy (as above), 1 read-only sweep of d_src (one sum
 * per workgroup written to d_dst, which must hold 8 x 4 x CUs bytes), 2
 * write-only fill of d_dst (d_src unused): the HBM ceilings of the three
 * access mixes. */
int tfg_stream_bench(void* d_dst, const void* d_src, long long bytes, int mode, int reps,
                     double* ms_avg);
/* No counterpart in the reference (measurement): the cost of the two ways of
 * separating dependent phases on this device, `grid` workgroups of 256
 * threads: *us_barrier = one grid-wide barrier inside a cooperative launch
 * (reps of them in one kernel), *us_boundary = one kernel boundary (reps
 * empty kernels back to back on one stream).  TFG_EINVAL when the grid
 * cannot be resident at once. */
int tfg_sync_bench(int grid, int reps, double* us_barrier, double* us_boundary);

/* SpMV engines: 0 LDS-staged rows (2040 nonzeros per block), 1 the same
 * with 4088, 2/3/4 sub-wavefront rows with 4/8/16 lanes per row, 5 the
 * 2040 stage with the next block's loads in flight, 6 the same reading a
 * 1-byte code per nonzero instead of its int32 column (col = row + table[code];
 * available when the matrix has at most 256 distinct col - row offsets, as
 * every structured grid has), 7 the same reading one byte per ROW: the id of
 * its offset sequence in a table of at most 64 sequences of at most 12
 * entries (a 7-point box grid has 27), so neither columns nor row pointers
 * are streamed.  8: symmetric matrices on a row pattern with at most 8
 * distinct positive offsets and no buffer cells: the values are held once,
 * as the diagonal and the upper diagonals (8 B per row and per upper entry
 * instead of 8 B per nonzero), refilled from val at the start of every solve
 * or product, which also checks that val is symmetric bit for bit; when it
 * is not, that call uses engine 7.  Engine 8 serves the CG / BiCG product
 * and tfg_spmv; every other kernel runs engine 7.  All give bit-identical
 * products; the engine
 * sets the SpMV launch grid and so the tree of the dot products fused into
 * it.  Default: engine 8 at 2M nonzeros and more, else 7, else 6, else 5
 * (what the matrix offers), so that a solve is bitwise
 * reproducible from run to run.  The autotuner is opt-in: tfg_spmv_tune
 * times every engine on d_val (us[9] receives each mean launch time, -1
 * for an engine the matrix or d_val does not offer) and
 * keeps the fastest; TFG_SPMV_AUTOTUNE=1 in the environment makes the first
 * product or solve of a handle do the same; tfg_spmv_select forces an engine
 * (-1: default choice again on next use). */
int tfg_spmv_tune(tfg_matrix* h, const double* d_val, int* kind, float* us);
int tfg_spmv_select(tfg_matrix* h, int kind);

/* Facts about the handle: sizes, chosen SpMV engine (-1: not chosen yet),
 * its grid, the number of 2040-nonzero row blocks, the number of distinct
 * col - row offsets when engine 6 is available (else 0), of distinct row
 * patterns when engine 7 is (else 0), of upper diagonals when engine 8
 * is (else 0), whether the last val engine 8 was set up for was
 * symmetric (1), not (0: that call used engine 7) or none was (-1), the
 * internal ordering (TFG_ORDER_*; the engine figures are then those of the
 * renumbered matrix), and the tiles of the ICC sweeps on a box grid (0: the
 * row-level sweeps).  Any pointer may be NULL. */
int tfg_matrix_info(tfg_matrix* h, int* n, int* nnz, int* spmv_kind,
                    int* spmv_grid, int* stage_blocks, int* offsets, int* patterns,
                    int* sym_diagonals, int* sym_state, int* ordering, int* icc_tiles);

/* ---- internal ordering ----------------------------------------------------
 * No counterpart in the reference.  A matrix whose numbering has no locality
 * (randomly numbered cells, as a permuted kNN or a badly ordered unstructured
 * grid) makes every gather of the SpMV a cache miss.  TFG_ORDER_RCM
 * renumbers rows and columns internally by reverse Cuthill-McKee (computed
 * once, on the host); every call then permutes val, b and x in and x (or
 * the product) out, so callers keep their own numbering.  Each row keeps
 * its entries in the caller's order, so products are bit-identical; only
 * the order of the global dot products changes, as between SpMV engines.
 * tfg_matrix_create applies it by itself to a matrix of 65536 rows or more
 * without row patterns whose numbering has no locality (mean |col - row| of
 * the off-diagonals above nt/64).  "icc" (an ordering-dependent
 * factorisation) and face assembly always run in the caller's order;
 * communication hooks drop the ordering; not with buffer cells
 * (TFG_ENOTSUP).  TFG_ORDER_NATURAL drops it. */
#define TFG_ORDER_NATURAL 0
#define TFG_ORDER_RCM 1
int tfg_matrix_reorder(tfg_matrix* h, int order);

/* ---- pressure-system assembly on the device ------------------------------
 * Builds A % val and b of the pressure-correction system from face data on
 * the device, so that a pressure solve copies neither to the GPU: the
 * matrix lines of the Rhie and Chow face loop (Rhie_And_Chow.f90:184-205),
 * the right-hand side from the volume fluxes and the boundary diagonal terms
 * (Compute_Pressure.f90:121-122, 147-169), the boundary regions' b
 * (:171-189) and, when `singular`, Set_Singular's a_ii * (1 + 1e-6)
 * (Solver_Mod/Set_Singular.f90:28-31).  Bit-identical to the serial loops.
 * Structure once per grid: c1, c2 = Grid % faces_c (1-based; c2 <= 0 on the
 * boundary), pos = A % pos (2 x n_faces, 1-based, pos(1,s) then pos(2,s);
 * ignored on the boundary), per boundary face diag_bc = 1 for a pressure or
 * ambient condition (Bnd_Cond_Type of c2) and b_region = the type of the
 * face's region: TFG_REGION_PRESSURE (b(c1) = 0, :173-178),
 * TFG_REGION_AMBIENT (b(c1) = fc(s) v_m(c1) pp % n(c2), :180-187) or
 * TFG_REGION_NONE.  The regions' assignments of b follow face order, which
 * is the reference's region loop when each region's faces form one range
 * and the ranges ascend with the region number (Faces_In_Region). */
#define TFG_REGION_NONE 0
#define TFG_REGION_PRESSURE 1
#define TFG_REGION_AMBIENT 2
typedef struct tfg_faces tfg_faces;
int tfg_faces_create(tfg_matrix* h, int n_faces, const int* c1, const int* c2,
                     const int* pos, const int* diag_bc, const int* b_region,
                     tfg_faces** faces);
/* The faces handle uses its matrix handle: assemble only while both exist.
 * Destroying it needs only itself, before or after the matrix handle. */
int tfg_faces_destroy(tfg_faces* faces);
/* Device pointers: fc = A % fc (n_faces), fw = Grid % f (n_faces), vm =
 * M % v_m (nt), vflux = v_flux % n (n_faces); val (nnz) and b (nt) out.
 * Stream-ordered on the matrix handle's stream (a following solve on the
 * same handle sees the result). */
int tfg_pressure_assemble(tfg_faces* faces, const double* fc, const double* fw,
                          const double* vm, const double* vflux, double* val,
                          double* b, int singular);
/* The same with pp = pp % n (device, pointing at cell 1; boundary cell c < 0
 * at pp[c - 1], Fortran's (-n_bnd:n) array passed as c_loc(pp(1))), which
 * ambient regions read; tfg_pressure_assemble (pp = NULL) returns
 * TFG_EINVAL when the faces have ambient regions. */
int tfg_pressure_assemble_ex(tfg_faces* faces, const double* fc, const double* fw,
                             const double* vm, const double* vflux, const double* pp,
                             double* val, double* b, int singular);
/* vel_max = max(MICRO, max over cells 1..nt of sqrt(u^2 + v^2 + w^2)), the
 * norm of the momentum solves before its Global % Max_Real
 * (Compute_Momentum.f90:201-204), from device u, v, w (nt each); the result
 * to host memory (waits for the handle's stream). */
int tfg_vel_max(tfg_matrix* h, const double* u, const double* v, const double* w,
                double* vel_max);
/* The smallest and largest value of a device vector over cells 1..nt: the
 * pressure range of Compute_Pressure.f90:104-109 (p_min, p_max before their
 * Global % Min/Max_Real); the pressure solve's norm is then
 * p_nor = max(p_max - p_min, p_nor_c, |p_drop_x|, |p_drop_y|, |p_drop_z|)
 * (:115-117).  Results to host memory (waits for the handle's stream). */
int tfg_range(tfg_matrix* h, const double* v, double* vmin, double* vmax);

/* ---- momentum assembly on the device ------------------------------------
 * Matrix, right-hand side and M % v_m of one velocity component from face
 * data, as Compute_Momentum.f90:119-366 builds them with Numerics_Mod
 * (Advection_Term with Min_Max and Advection_Scheme, Inertial_Term,
 * Under_Relax): advection in the chosen scheme with the implicit upwind
 * part when blend_matrix is set, the diffusion face loop with the caller's
 * face viscosity and stress (Turb % Face_Vis / Face_Stress), explicit cross
 * diffusion, inertia, global pressure drop, cell forces, the pressure
 * gradient, optional surface tension, v_m = vol / a_ii and under-
 * relaxation.  Bit-identical to the serial loops.  Not included: user
 * forces (User_Mod_Force) and PISO's explicit part; vel_max is tfg_vel_max.
 * Needs the faces handle of tfg_faces_create (its pos, dia and c2 < 0 for
 * boundary cells; each boundary cell on one face, each matrix entry set by
 * one face).  All pointers are device pointers.  Cell arrays point at cell
 * 1; those marked [b] also hold the boundary cells c = -n_bnd..-1 at offset
 * c - 1 (Fortran's (-n_bnd:n) array passed as c_loc(a(1))). */
#define TFG_ADV_CENTRAL 0
#define TFG_ADV_UPWIND 1
#define TFG_ADV_LUDS 2
#define TFG_ADV_QUICK 3
#define TFG_ADV_SMART 4
#define TFG_ADV_GAMMA 5
#define TFG_ADV_MINMOD 6
#define TFG_ADV_BLENDED 7
#define TFG_ADV_SUPERBEE 8
#define TFG_ADV_AVL_SMART 9
#define TFG_TD_LINEAR 0      /* backward Euler */
#define TFG_TD_PARABOLIC 1   /* BDF2 */
#define TFG_TD_NONE 2        /* no inertial term */
typedef struct {
  int component;      /* 1, 2, 3: u, v, w (which of s, d, grad are i, j, k) */
  int n_bnd;          /* boundary cells */
  int adv_scheme;     /* TFG_ADV_* (ui % adv_scheme) */
  int blend_matrix;   /* ui % blend_matrix */
  int td_scheme;      /* TFG_TD_* (ui % td_scheme) */
  double blend;       /* ui % blends(1), BLENDED only */
  double dt, urf, p_drop;  /* Flow % dt, ui % urf, bulk % p_drop_i */
  /* per face */
  const double *f, *fw, *fc, *v_flux, *vis_eff, *f_stress;  /* f_stress may be NULL */
  const double *sx, *sy, *sz, *dx, *dy, *dz;
  const int* wall;    /* 1: boundary face with INFLOW, WALL, CONVECT or WALLFL */
  /* per cell: phi_n [b] = ui % n, phi_o, phi_oo, gx gy gz [b] = ui % x y z,
   * uj_i uk_i [b] = the other components' gradients along i, density [b],
   * vol, p_i = p % x|y|z, cell_f (may be NULL) = Flow % cell_fi, st (may be
   * NULL) = Vof % surf_fi */
  const double *phi_n, *phi_o, *phi_oo, *gx, *gy, *gz, *uj_i, *uk_i, *density, *vol, *p_i,
      *cell_f, *st;
} tfg_momentum;
/* val (nnz), b (nt), v_m (nt) out; stream-ordered like tfg_pressure_assemble. */
int tfg_momentum_assemble(tfg_faces* faces, const tfg_momentum* m, double* val, double* b,
                          double* v_m);

/* ---- Rhie-Chow face fluxes on the device ----------------------------------
 * The volume flux of every domain face (c2 > 0), Rhie_And_Chow.f90:86-259:
 * the interpolated cell velocities through the face plus the pressure
 * smoothing a12 (p1 - p2) + fc (grad p v_m)_f . d, with the optional VOF
 * surface-tension term, Choi's time-step correction and Gu's body-force
 * correction.  Each face on its own: bit-identical to the serial loop.
 * Boundary faces are not written.  The matrix lines of the same loop are
 * tfg_pressure_assemble's.  Device pointers; cell arrays hold the cells
 * 1..n (Cells_In_Domain_And_Buffers), face arrays n_faces entries. */
typedef struct {
  /* per face: Grid % f, A % fc, Grid % sx sy sz, Grid % dx dy dz */
  const double *f, *fc, *sx, *sy, *sz, *dx, *dy, *dz;
  /* per cell: u % n, v % n, w % n, p % n, p % x y z, M % v_m */
  const double *u, *v, *w, *p, *px, *py, *pz, *v_m;
  /* Choi's correction (choi != 0, applied when the caller's
   * Time % Curr_Dt() - Time % First_Dt() > 3 holds): u/v/w % o and oo,
   * density, v_flux % o and oo, dt, and w_o, w_oo (1, 0 for LINEAR, 2,
   * -0.5 for PARABOLIC) */
  int choi;
  const double *u_o, *u_oo, *v_o, *v_oo, *w_o_, *w_oo_, *density, *flux_o, *flux_oo;
  double dt, w_o, w_oo;
  /* Gu's correction (gu != 0): Flow % cell_fx fy fz (cells), face_fx fy fz */
  int gu;
  const double *cell_fx, *cell_fy, *cell_fz, *face_fx, *face_fy, *face_fz;
  /* surface tension (vof != 0 and sigma > TINY = 1e-15): Vof % curv, fun % n, fun % x y z */
  int vof;
  double sigma;
  const double *curv, *col, *col_x, *col_y, *col_z;
} tfg_rhie_chow;
/* v_flux (n_faces) out on the domain faces; stream-ordered. */
int tfg_rhie_chow_fluxes(tfg_faces* faces, const tfg_rhie_chow* rc, double* v_flux);

/* ---- several MPI ranks -----------------------------------------------------
 * A matrix with buffer cells (ni < nt: rows ni+1..nt are copies of other
 * ranks' cells) is solved with the caller's communication, which stands in
 * for Comm_Mod where Cg.f90 calls it:
 *   allreduce(ctx, values, ops, count): combine `count` host doubles in place
 *     over all ranks, ops[i] = TFG_SUM | TFG_MIN | TFG_MAX
 *     (Global % Sum_Real / Min_Real / Max_Real,
 *     Comm_Mod/Parallel/Global/Sum_Real.f90 and its siblings);
 *   exchange(ctx, d_x): refresh the buffer cells d_x[ni..nt-1] of a device
 *     vector from their owners (Grid % Exchange_Cells_Real,
 *     Grid_Mod/Parallel/Exchange_Cells_Real.f90:2-129); d_x[0..ni-1] is current.
 * Both return 0 on success; anything else aborts the solve (TFG_EINVAL).
 * Called from the thread that runs the solve, with the device idle.  Every
 * rank makes the same calls in the same order.  FP64 cg / bicg with jacobi,
 * icc or none; "icc" is the reference's per-rank DIC over rows 1..ni
 * (buffer columns left out, Prec_Solve.f90:89), so it needs each row's
 * diagonal first and its other columns ascending in the rank's own
 * numbering.  x's buffer cells must be current on entry, as in the
 * reference, and are refreshed at exit (Cg.f90:240). */
#define TFG_SUM 0
#define TFG_MIN 1
#define TFG_MAX 2
typedef int (*tfg_allreduce_fn)(void* ctx, double* values, const int* ops, int count);
typedef int (*tfg_exchange_fn)(void* ctx, double* d_x);
int tfg_matrix_set_comm(tfg_matrix* h, tfg_allreduce_fn allreduce,
                        tfg_exchange_fn exchange, void* ctx);

/* ---- batched solves ------------------------------------------------------
 * No counterpart in the reference, which solves one system per Solve_Native
 * call: nsys independent systems solved together, each with exactly the
 * semantics of tfg_solve_native_ex (its own normalisation, stopping test,
 * niter and fin_res).  Systems that stop are masked and dropped from the
 * launches; a chunk of iterations is replayed as a hipGraph (opt->path =
 * TFG_PATH_KERNELS launches kernels one by one instead).  FP64, "cg"/"bicg",
 * jacobi or none.  A system whose entries lie on its diagonal and at most 3
 * upper diagonals and their mirrors (a hex grid), rows stored diagonal first
 * then by ascending column, takes its products from zero-padded diagonals
 * when its val is symmetric bit for bit (checked per solve), with the CSR
 * product's bits; TFG_BATCH_SPMV=csr in the environment keeps every system
 * on CSR. */
typedef struct tfg_batch tfg_batch;

/* n[s], nnz[s]: rows and nonzeros of system s; row, col, dia: the systems'
 * 1-based Matrix_Type arrays concatenated system after system (row holds
 * n[s]+1 entries per system). */
int tfg_batch_create(int nsys, const int* n, const int* nnz, const int* row,
                     const int* col, const int* dia, tfg_batch** handle);
int tfg_batch_destroy(tfg_batch* handle);

/* val, x, b: concatenated like the structure (host or device pointers);
 * niter[nsys], fin_res[nsys] out; norm: NULL or nsys values.  rep (may be
 * NULL): iterations = the largest niter, solve_ms = device time of the batch.
 * opt->precision TFG_FP32_IR runs every system in the FP32-sweep form of
 * tfg_solve_native_ex (niter = FP32 iterations, fin_res = the measure of the
 * last FP64 true residual), jacobi or none. */
int tfg_batch_solve(tfg_batch* h, const char* solver, int solver_len,
                    const char* prec, int prec_len, const double* val, double* x,
                    double* b, int miter, int* niter, double tol, double* fin_res,
                    const double* norm, int device_ptrs, const tfg_options* opt,
                    tfg_report* rep);

/* ms[nsys]: time from the start of the last batched solve to the stop of
 * each system (device clock). */
int tfg_batch_times(tfg_batch* h, double* ms);
int tfg_batch_info(tfg_batch* h, int* nsys, int* n, int* nnz, int* blocks);

/* The batch autotuner (no counterpart in the reference).  On the first
 * solve of a handle in each precision, unless TFG_BATCH_TUNE=0 in the
 * environment or miter < 32, short timed probes (the first 32 iterations
 * from the caller's x0, restored afterwards) choose the iterations each
 * replayed graph holds (FP64: 8, 16 or 32; the FP32 form keeps 8, its
 * refinement cadence) and, per system-size class (bit 0: < 32768 rows,
 * bit 1: < 262144, bit 2: larger), whether systems on diagonals take their
 * products from the diagonals (bit set) or the CSR rows, and the rows each
 * thread of the block kernels (p update, x and r update, diagonal
 * products) takes: 1, or 2 (one from each of two row blocks per
 * workgroup).  Every choice gives the
 * same bits.  Reports the choice for precision (TFG_FP64 | TFG_FP32_IR):
 * chunk, diag_classes and rows_per_thread -1 before the first solve;
 * probes timed and the best probe's device time.  Any pointer may be
 * NULL.  TFG_BATCH_ROWS=1|2 in the environment forces the rows per thread. */
int tfg_batch_tuning(tfg_batch* h, int precision, int* chunk, int* diag_classes,
                     int* rows_per_thread, int* probes, double* probe_ms);

/* ---- allocation-free solve ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory during a solve (code inside HIP graph capture of their own, tight
 * serving loops, harnesses that forbid hipMalloc): tfg_solve_native_ex with
 * device pointers whose work vectors (r, p, q, the scaled diagonal), partial
 * sums, solver state and the symmetric engine's diagonals live in one static
 * device array of TFG_NOALLOC_MAX_BYTES per GPU, held from the library's
 * first tfg_matrix_create on that GPU until the process ends.  The call
 * makes no hipMalloc / hipFree and no other allocating call; x is iterated in
 * place.  Same kernels, SpMV engine, launch shapes and so the same bits as
 * tfg_solve_native_ex's FP64 path.  Scope: "cg" / "bicg" (the reference's),
 * jacobi or none, FP64, no buffer cells, no internal ordering, no timing
 * (TFG_ENOTSUP otherwise); a system needing more scratch than the maximum
 * (about 32 bytes per row, plus 8 (K + 1) per row with the symmetric engine:
 * roughly 0.75M rows of a 7-point grid) returns TFG_EINVAL with the bytes it
 * needs.  Every call on one GPU shares that GPU's array: calls that overlap
 * in time on one GPU (two threads, two streams) race and must be serialised
 * by the caller. */
#define TFG_NOALLOC_MAX_BYTES (48LL << 20)
int tfg_solve_native_noalloc(tfg_matrix* h, const char* solver, int solver_len,
                             const char* prec, int prec_len, const double* d_val,
                             double* d_x, const double* d_b, int miter, int* niter,
                             double tol, double* fin_res, const double* norm,
                             const tfg_options* opt, tfg_report* rep);

const char* tfg_last_error(void);
int tfg_version(void);

#ifdef __cplusplus
}
#endif
#endif
