// Internal declarations shared by the tflows_gpu kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tflows_gpu.h"

namespace tfg {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kRowsPerBlock = 256; // SpMV rows per staged block (one per thread)
// Nonzeros per staged block of the stream engines: 2040 (4088) fit 4 (8)
// 16-byte chunks per thread with the chunk of misalignment at each end.
constexpr int kCap = 2040;
constexpr int kCap2 = 4088;
// Row-pattern engine: at most kPatCount distinct rows of at most kPatMax
// entries.
constexpr int kPatCount = 64;
constexpr int kPatMax = 12;
// Distinct positive column offsets the symmetric engine stores (upper
// diagonals; a 7-point hex grid has 3), and the most it takes two rows per
// thread with.
constexpr int kSymMax = 8;
constexpr int kSymPairMax = 4;
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dealt round robin
constexpr double kHuge = 1.0e+15;  // Const_Mod.f90:84-93 (HUGE, TINY)
constexpr double kTiny = 1.0e-15;

enum Prec { PREC_NONE = 0, PREC_JACOBI = 1, PREC_ICC = 2 };

// Reduction operators of the partial sums (same codes as TFG_SUM/MIN/MAX).
enum RedOp { OP_SUM = 0, OP_MIN = 1, OP_MAX = 2 };

// Scalars of one solve, resident in device memory.  Each field is written by
// one thread (workgroup 0 of a kernel, or a one-workgroup kernel) and read by
// later kernels; fields read inside the kernel that writes them are never
// the ones it writes (rho is double-buffered by iteration parity).
struct State {
  double fn;        // 1 / mean(diag)                     (Cg.f90:62-76)
  double rho[2];    // rho_k = (z, r) in rho[k & 1]       (Cg.f90:150-156)
  double res;       // last stopping measure
  double bnorm;     // ||b_scaled||_2
  double norm;      // optional normalisation (valid when has_norm)
  double tol;
  double aux0;      // scratch results (true residual, refinement norms)
  double aux1;
  double aux2;
  int has_norm;
  int criterion;
  int done;         // 1: stop, every later phase returns at once
  int niter;        // iteration at which done was raised
  int status;       // TFG_OK or TFG_EBREAKDOWN
  int iter;         // last iteration whose stopping test was evaluated
  double alpha;     // CGS / BiCGSTAB, deferred-x CG: alpha of the current iteration
  double omega;     // BiCGSTAB: omega of the last iteration
  int fold;         // deferred-x CG: x += alpha p of the last update still to do
  double beta;      // multi-rank CG: beta from the global reduction
  int comm;         // multi-rank CG: fn is given (global), not computed
  int kcur;         // graph path: the iteration the next kernels run
  // FP32 sweeps in FP64 refinement (tfg_ir.hip)
  double defl_e;    // w'Aw of the deflation vector w = 1 (0: no deflation)
  double rref;      // ||r32||^2 at the last restart
  double xmin, xmax;  // x range at the last refinement (NRMS estimate)
  double delta;     // refinement when ||r32|| fell by this since the restart
  int refine;       // 1: the refinement kernels of this chunk end run
  int kref;         // iteration of the last restart (0: the start)
  int miter;        // iteration cap
  int force;        // > 0: a refinement at least every `force` iterations
  int defl_mode;    // 0 auto, 1 off, 2 on (TFG_DEFLATE_*)
  int sym_vals;     // the values are symmetric (deflation allowed); cleared by
                    // any thread of the check that finds an asymmetric pair
  int nrefine;      // refinements (true residuals) so far
  // batched FP32 sweeps (tfg_batch.hip): per solve sum_i s_i, sum_i a_ii fn,
  // sum_i b_i fn as a double-double (hi, lo); the coarse correction c of
  // the pending restart
  double dsum[4];
  double corr;
};

// Partial sums written by one kernel and swept by the next ones.
struct Partials {
  double* diag;     // [1][pstride]  sum of the diagonal
  double* pro;      // [7][pstride]  prologue sums
  double* spmv;     // [1][pstride]  (p, q)
  double* upd;      // [5][pstride]  NRMS sum, x min, x max, (z,r), (r,r)
  double* aux;      // [2][pstride]  true-residual sums
  double* ext;      // [6][pstride]  BiCGSTAB: sums of s, then (t,s), (t,t)
  double* irc;      // [4][pstride]  FP32 refinement form: conversion sums
  int pstride;
  int g_vec;        // grid of the element-wise kernels (producers of diag)
  int g_spmv;       // grid of the SpMV kernels (producers of pro/spmv/aux)
  int g_upd;        // grid of the kernel that produces upd (method dependent)
};

// Symmetric engine geometry (by value in the kernel arguments).
struct SymDiag {
  long long stride;   // slot stride of sd, even
  int k;              // upper diagonals, 1..kSymMax
  int pair;           // 1: two rows per thread (d_0 = 1, other d_k even, n even,
                      //    k <= kSymPairMax; blocks of 512 rows at even starts)
  int off[kSymMax];   // d_0 < d_1 < ... < d_{k-1}
};

// Everything a CG phase kernel needs.  Device pointers.
struct CgArgs {
  int n;               // rows solved (ni)
  const int* row;      // 0-based row starts, n+1
  const int* col;      // 0-based columns, nnz
  const int* blk;      // SpMV block row boundaries, nblk+1
  int nblk;
  const uint8_t* code; // offset codes (coded engine) or null
  const int* tab;
  int ntab;
  const uint8_t* pat;  // row patterns (pattern engine) or null
  const int* poff;
  const int* plen;
  const int* pslot;    // symmetric engine: value slot of each pattern entry
  const int* sorder;   // symmetric engine: (start, end) row blocks in processing order
  int nsorder;
  const double* sd;    // symmetric engine: diagonal and upper diagonals
  SymDiag sdg;         // symmetric engine: offsets, slot stride, row pairing
  const double* val;   // unscaled values (caller's)
  const double* b;     // unscaled right-hand side (caller's)
  double* x;
  double* r;
  double* p;
  double* q;
  double* ad;          // scaled diagonal a_ii * fn
  Partials pt;
  State* st;
};

// CSR structure as the row engines see it.
struct SpmvShape {
  int n;
  const int* __restrict__ row;   // 0-based, n+1
  const int* __restrict__ col;   // 0-based
  const int* __restrict__ blk;   // row-block boundaries (stage engines)
  int nblk;
  const uint8_t* __restrict__ code;  // offset code of each nonzero (coded engine)
  const int* __restrict__ tab;       // offset table: col = row + tab[code]
  int ntab;
  const uint8_t* __restrict__ pat;   // row pattern of each row (pattern engine)
  const int* __restrict__ poff;      // kPatCount x kPatMax column offsets
  const int* __restrict__ plen;      // kPatCount pattern lengths
  // symmetric engine: sd holds the diagonal (sd[i]) and the upper diagonals
  // (sd[(k + 1) S + i] = a(i, i + d_k), 0 < d_0 < d_1 < ..., zero where row
  // i has no such entry; S = sdg.stride leaves the largest d_k + 2 zeros after
  // every slot); pslot gives per pattern entry its slot: 0 diagonal, k + 1
  // the row's own upper slot k, -(k + 1) a lower entry (slot k of row
  // i + offset), for the set-up kernel
  const int* __restrict__ pslot;
  const double* __restrict__ sd;
  const int* __restrict__ sorder;    // symmetric engine: (start, end) row blocks
  int nsorder;                       //   in processing order, and their count
  SymDiag sdg;
};

// Textbook methods for non-symmetric systems (tfg_nonsym.hip).
enum NsMethod { NS_NONE = 0, NS_BICG = 1, NS_CGS = 2, NS_BCGS = 3 };

// Extra vectors of those methods; roles per method are listed in
// tfg_nonsym.hip.
struct NsArgs {
  const double* valt;  // transposed values (BiCG), unscaled
  double* rt;          // shadow residual
  double* v1;
  double* v2;
  double* v3;
  double* v4;
};

// SpMV row engines (see tfg_spmv.h).
enum SpmvKind { K_STAGE2048 = 0, K_STAGE4096 = 1, K_WAVE4 = 2, K_WAVE8 = 3, K_WAVE16 = 4,
                K_PIPE2048 = 5, K_PIPE2048_CODED = 6, K_PIPE_PATTERN = 7, K_SYM_DIA = 8,
                kKinds = 9 };
// FP32 sweeps of iterative refinement (tfg_ir.hip): FP32 vectors and
// matrix in the row-excess form, FP64 scalars, reductions, solution and
// true residual.  Partial sums go to the FP64 solve's buffers.
struct IrArgs {
  int n;
  float* val32;         // CSR engines: fl32(a_ij fn), the diagonal slot s_i
  float* sd32;          // symmetric engine: slot 0 s_i, slot k+1 fl32(a(i, i+d_k) fn)
  long long stride32;   // slot stride of sd32 (floats)
  const int* sorder32;  // symmetric engine: (start, end) blocks of 4-row groups
  int nsorder32;
  int quad;             // symmetric engine: four rows per thread
  double* s64;          // s_i = sum_j a_ij fn (compensated, CSR order)
  float* s32;           // fl32(s_i)
  float* rd32;          // fl32(1 / (a_ii fn))
  float* r32;
  float* p32;
  float* q32;
  float* e32;           // FP32 correction since the last refinement
  int prec;             // PREC_JACOBI or PREC_NONE
};

// Level-scheduled DIC (tfg_icc.hip).  Device arrays except the two level
// offset tables, which the host reads to size each level's launch.
// DIC sweeps of a box grid (tfg_icc.hip): rows (x, y, z), x slowest, with
// every entry a box neighbour (offsets 1, nz, ny*nz).  The values sit in
// seven zero-padded diagonals; the sweeps run over tiles of at most
// kBoxTX x kBoxTY x kBoxTZ cells (z is the fastest axis), one workgroup
// per tile (one thread per cell) and one launch per tile level.  Measured at
// 200^3 (ICC-CG solve): cubes of side 4, 6, 7, 8, 9, 10, 12, 16 ran 1.89,
// 1.59, 1.52, 1.34, 1.43, 1.35, 1.57, 1.78 s (smaller tiles take more
// launches, larger ones more barriers); 8x8x16 1.32, 16x8x8 1.33, 8x8x12
// 1.36, 8x16x16 1.47 s.  The level loop runs on one wave, a lane per (x, y)
// pencil, so kBoxTX * kBoxTY is 64.
constexpr int kBoxTX = 8, kBoxTY = 8, kBoxTZ = 16;
constexpr int kBoxThreads = 1024;
struct BoxIcc {
  int nx, ny, nz;
  int tx, ty, tz;           // tile sides (<= kBoxTX, kBoxTY, kBoxTZ)
  int ntx, nty, ntz;        // tiles per side
  const int* tiles;         // tile ids X * nty * ntz + Y * ntz + Z, sorted by X + Y + Z
  double* dg;               // a_ii
  double* lo[3];            // a(i, i - off) for off = ny*nz, nz, 1 (zero when absent)
  double* up[3];            // a(i, i + off), same order
};

struct IccArgs {
  int n;
  const int* row;       // 0-based
  const int* col;
  const int* rows_f;    // rows sorted by forward level
  const int* rows_b;    // rows sorted by backward level
  double* d;            // 1 / d_inv (the reference's d)
  double* dinv;         // a_ii - sum d_k a_ik^2 (the reference's d_inv)
  double* y;            // forward-substitution result
  int nlev_f, nlev_b;
  const int* off_f_host;  // host: nlev_f + 1 level starts in rows_f
  const int* off_b_host;  // host: nlev_b + 1 level starts in rows_b
  int box;                // 1: the tiled sweeps of a box grid
  BoxIcc bx;
  const int* toff_host;   // host: tile-level starts in bx.tiles (ntx + nty + ntz - 1 levels)
};

// Pressure-system assembly from face data (tfg_assembly.hip).
struct AsmArgs {
  int n;                 // cells (rows) assembled: nt
  const int* row;        // 0-based CSR row starts (diagonal first)
  const int* cf_off;     // n+1: each cell's faces in cf_face
  const int* cf_face;    // face index * 2 + (1 if the cell is the face's c2)
  const int* ent_face;   // nnz: last face setting each off-diagonal entry, -1 none
  const int* c1;         // n_faces, 0-based
  const int* c2;         // n_faces, 0-based; -1 for a boundary face
  const uint8_t* flags;  // n_faces: 1 boundary diagonal term, 2 b(c1) = 0 (pressure
                         // region), 4 b(c1) = fc v_m(c1) pp(c2) (ambient region)
  const int* c2b;        // n_faces: c2 as given (boundary cells < 0)
  const double* pp;      // pp % n at cell 1 (boundary cell c at pp[c - 1]); ambient only
  const double* fc;
  const double* fw;
  const double* vm;
  const double* vflux;
  double* val;
  double* b;
  int singular;
};

// Momentum assembly (tfg_assembly.hip).  Cell arrays point at cell 1 (a
// 1-based cell C is element C - 1; boundary cells C < 0 below it).
struct MomArgs {
  int n, nb;
  const int* row;
  const int* cf_off;
  const int* cf_face;
  const int* ent_face;
  const int* c1;         // 0-based
  const int* c2b;        // 1-based as given (boundary cell < 0)
  int comp, adv, blend_matrix, td;
  double blend, dt, urf, p_drop;
  const double *f, *fw, *fc, *v_flux, *vis_eff, *f_stress;
  const double *sv[3], *dv[3];
  const int* wall;
  const double *phi, *phi_o, *phi_oo, *g[3], *uj_i, *uk_i, *rho, *vol, *p_i, *cell_f, *st;
  double* pmin;          // n scratch values: min / max of phi over each cell's faces
  double* pmax;
  double *val, *b, *v_m;
};
hipError_t launch_momentum(hipStream_t stream, const MomArgs& m);

// Rhie-Chow face fluxes (tfg_assembly.hip); c1 0-based, c2 0-based or -1.
struct RcArgs {
  int n_faces;
  const int* c1;
  const int* c2;
  tfg_rhie_chow r;
  int vof;               // surface tension term active
  double* v_flux;
};
hipError_t launch_rhie_chow(hipStream_t stream, const RcArgs& a);

// dst[k] = src[map[k]], k < count (permutation of values and vectors)
hipError_t launch_gather(hipStream_t stream, const double* src, const int* map, double* dst,
                         long long count);

// Launch shapes chosen per matrix.
struct Launch {
  hipStream_t stream;
  int grid_vec;    // workgroups of the element-wise kernels
  int grid_spmv;   // workgroups of the SpMV kernels (resident, multiple of kXcds)
  int kind;        // SpmvKind
  bool vec_val;    // val is 16-byte aligned (wide nontemporal loads)
  bool sym;        // products by the symmetric engine (val checked symmetric)
};

// Resident SpMV workgroups per CU for an engine (occupancy query).
int spmv_blocks_per_cu(int kind);

hipError_t launch_diag_sum(const Launch& L, const CgArgs& a);
hipError_t launch_prologue(const Launch& L, const CgArgs& a, int prec);
hipError_t launch_start(const Launch& L, const CgArgs& a, int rho_slot = 6);
hipError_t launch_pupdate(const Launch& L, const CgArgs& a, int prec, int k, bool dx = false);
hipError_t launch_fold_x(const Launch& L, const CgArgs& a);
hipError_t launch_spmv_pq(const Launch& L, const CgArgs& a, int k);
// Symmetric engine set-up for one val: sd from val, and *flag = 1 when some
// lower entry is not bitwise equal to its mirror (then sd must not be used).
hipError_t launch_sym_pack(hipStream_t stream, int grid, const SpmvShape& m,
                           const double* val, double* sd, int* flag);
hipError_t launch_update(const Launch& L, const CgArgs& a, int prec, int k, bool dx = false);
hipError_t launch_finish(const Launch& L, const CgArgs& a, int k, int check_omega = 0);

// DIC preconditioner (tfg_icc.hip).  st = null: unscaled matrix (fn = 1)
// and no stop flag.
hipError_t launch_icc_form(hipStream_t s, const IccArgs& c, const double* val, const State* st);
hipError_t launch_icc_apply(hipStream_t s, const IccArgs& c, const double* val, const State* st,
                            const double* r, double* z);
hipError_t launch_icc_check(const Launch& L, const CgArgs& a, int k);
hipError_t launch_icc_rho(const Launch& L, const CgArgs& a, const double* z);
hipError_t launch_icc_pupdate(const Launch& L, const CgArgs& a, const double* z, int k);
hipError_t launch_jacobi_apply(hipStream_t s, int n, const int* row, const double* val,
                               const double* r, double* z);

hipError_t launch_assemble(hipStream_t stream, const AsmArgs& g);
hipError_t launch_range(hipStream_t stream, int n, const double* v, double* part, int G,
                        double* out);
hipError_t launch_vel_max(hipStream_t stream, int n, const double* u, const double* v,
                          const double* w, double* part, int G, double* out);

// Multi-rank CG (tfg_comm.hip).
hipError_t launch_cm_local(hipStream_t s, const double* part, int pstride, int G, int nv,
                           const int* ops, double* out);
hipError_t launch_cm_pupdate(const Launch& L, const CgArgs& a, int prec, const double* z, int k);
hipError_t launch_cm_update(const Launch& L, const CgArgs& a, int prec);

// Textbook methods (tfg_nonsym.hip).
hipError_t launch_ns_init(const Launch& L, const CgArgs& a, const NsArgs& s);
hipError_t launch_transpose_pos(hipStream_t stream, int n, const int* row, const int* col,
                                int* tpos, int* bad);
hipError_t launch_gather_valt(const Launch& L, long long nnz, const int* tpos, const double* val,
                              double* valt);
// One iteration k of method m; ev0/ev1 (may be null) bracket its first product.
hipError_t launch_ns_iteration(const Launch& L, const CgArgs& a, const NsArgs& s, int method,
                               int prec, int k, hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_true_residual(const Launch& L, const CgArgs& a);

hipError_t launch_spmv_plain(const Launch& L, const SpmvShape& m, const double* val,
                             const double* p, double* q);
hipError_t launch_copy16(hipStream_t stream, int grid, const void* src, void* dst, long long n16);
// out: one double per workgroup (grid entries)
hipError_t launch_read16(hipStream_t stream, int grid, const void* src, double* out, long long n16);
hipError_t launch_fill16(hipStream_t stream, int grid, void* dst, long long n16);
hipError_t launch_zero16(hipStream_t stream, void* dst, long long n16);
hipError_t launch_grid_barriers(hipStream_t stream, int grid, int reps, double* sink);
hipError_t launch_boundary(hipStream_t stream, int grid, double* sink, int r);
int grid_barrier_fit(int* per_cu);

// FP32 sweeps in FP64 refinement (tfg_ir.hip).  L: the FP64 engine (the
// true residual); L32: the FP32 product's engine (kind 8 = the symmetric
// FP32 engine on f.sd32, else a CSR engine on f.val32).
hipError_t launch_ir_convert(const Launch& L, const Launch& L32, const CgArgs& a,
                             const IrArgs& f, int nnz);
// State::sym_vals = 0 when some val[j] differs from val[tpos[j]] bitwise.
hipError_t launch_ir_symcheck(hipStream_t s, int grid, const double* val, const int* tpos,
                              long long nnz, State* st);
hipError_t launch_ir_widen(const Launch& L, const CgArgs& a, const float* r32, double* r64);
hipError_t launch_ir_icc_rho(const Launch& L, const CgArgs& a, const IrArgs& f, const double* z64,
                             float* z32);
hipError_t launch_ir_pupdate(const Launch& L, const CgArgs& a, const IrArgs& f);
hipError_t launch_ir_spmv(const Launch& L32, const CgArgs& a, const IrArgs& f);
hipError_t launch_ir_update(const Launch& L, const CgArgs& a, const IrArgs& f, bool defl,
                            bool nrms);
// Chunk end: fold (when a refinement is due), FP64 true residual, stopping
// test and restart; launch_ir_restart is the last two alone (the start).
hipError_t launch_ir_refine(const Launch& L, const CgArgs& a, const IrArgs& f);
hipError_t launch_ir_restart(const Launch& L, const CgArgs& a, const IrArgs& f);
// x += e after a breakdown.
hipError_t launch_ir_fold(const Launch& L, const CgArgs& a, const IrArgs& f);
// The FP64 true residual of a refinement (when State::refine) with its
// stopping sums and (s, x) (pt.ext).
hipError_t launch_ir_resid(const Launch& L, const CgArgs& a, const IrArgs& f);

// FP32 sweeps of the textbook methods (tfg_ir_ns.hip); roles of the vectors
// per method are listed there.
struct NsArgs32 {
  const float* valt;   // BiCG: the FP32 transpose in the row-excess form
  float* rt;           // shadow residual
  float* v1;
  float* v2;
  float* v3;
  float* v4;
};
hipError_t launch_ns32_valt(const Launch& L, const CgArgs& a, const int* tpos, float* valt);
// inner restart from the FP64 true residual a.r (inner State `in`)
hipError_t launch_ns32_restart(const Launch& L, const CgArgs& a, const IrArgs& f,
                               const NsArgs32& s, State* in, double tol_in, int method);
hipError_t launch_ns32_iteration(const Launch& L, const Launch& L32, const CgArgs& a,
                                 const IrArgs& f, const NsArgs32& s, State* in, int method,
                                 int k, hipEvent_t ev0, hipEvent_t ev1);
// x += e, FP64 true residual, stopping test (niter = iters; at miter stop)
hipError_t launch_ns32_fold_test(const Launch& L, const CgArgs& a, const IrArgs& f, int iters,
                                 int miter);
// The FP32 row-excess product y = s_i x_i + sum_{j != i} a_ij (x_j - x_i)
// of an FP32 matrix already converted (building block, tfg_spmv_fp32).
hipError_t launch_spmv_lap32(const Launch& L32, const SpmvShape& m, const IrArgs& f,
                             const float* x, float* y);
hipError_t launch_residual_plain(const Launch& L, const SpmvShape& m, const double* val,
                                 const double* b, const double* x, double* r);

}  // namespace tfg
