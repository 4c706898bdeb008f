// Batched FP64 CG: many independent Solve_Native calls in one pass per
// phase.  No counterpart in the reference (which solves one system per
// call); each system gets exactly the single-system semantics (its own
// 1/mean(diag) normalisation, stopping test and niter / fin_res conventions,
// Cg.f90:62-253), so a batch equals that many Solve_Native calls.
//
// Layout: the systems' CSR arrays are concatenated into one block-diagonal
// matrix (global 0-based rows and columns).  Row blocks (<= 256 rows,
// <= kCap nonzeros) never straddle two systems.  Every block-level kernel
// runs one workgroup per block (the vector kernels: one or two blocks) of
// the ACTIVE list (blocks of systems not yet stopped) and publishes one
// partial per block; a one-workgroup-per-system kernel then sweeps that
// system's partials in a fixed order and writes its scalars.  A system that
// stops is masked at once (its workgroups return) and removed from the
// active list at the next compaction, so late iterations only launch the
// systems still running.  The products run as two kernels over two lists:
// the systems on diagonals (SysDiag) and the others (CSR rows).  Per
// iteration:
//   pupdate (blocks) | spmv + (p,q) (blocks: diagonals, CSR) | alpha
//   (systems) | update + sums (blocks) | check (systems: stop test, beta)
// The kernels read their iteration number from the system's state, so a
// chunk of iterations is captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "tfg_host.h"
#include "tfg_spmv.h"

namespace tfg {

namespace {

// Partial rows of the batch (each nblk long).
// FP32 form: BP_UPD + 0..3 (z, r), (s, z), (r, r), sum (r / a_ii)^2 of the
// FP32 residual; BP_IRC + 0..3 sum s_i, sum a_ii fn, sum b_i fn (hi, lo);
// BP_RES + 0..4 the FP64 true residual's NRMS sum, x min, x max, (r, r),
// (s, x).
enum { BP_DIAG = 0, BP_PRO = 1, BP_SPMV = 8, BP_UPD = 9, BP_IRC = 14, BP_RES = 18, BP_ROWS = 23 };

// FP32 sweeps in FP64 refinement: refinement decisions every kBt32Check
// iterations (and at the cap), as the single-system form takes them.
constexpr int kBt32Check = 8;

// FP32 arrays of the refinement form (allocated on the first FP32 solve).
struct Bt32 {
  float* val32;   // every system: fl32(a_ij fn), the diagonal slot s_i (row-excess form)
  float* sd32;    // systems on diagonals: slot 0 s_i, slot t + 1 upper diagonal t (SysDiag layout)
  double* s64;    // s_i = sum_j a_ij fn, compensated, CSR order
  float* s32;     // fl32(s_i)
  float* rd32;    // fl32(1 / (a_ii fn))
  float* r32;
  float* p32;
  float* q32;
  float* e32;     // FP32 correction since the last refinement
  int* asym;      // per system: 1 when its values are not symmetric bit for bit
};

// A system whose entries lie on its diagonal and k <= kBatchSym upper
// diagonals and their mirrors (a hex grid: k = 3), each row stored diagonal
// first and then by ascending column.  Its products read the values once,
// from sd + base: the diagonal, then upper diagonal j at (j + 1) * stride,
// zero where a row lacks the entry, each followed by at least off[k-1] + 2
// zeros (so a lower entry of an early row reads zeros).
constexpr int kBatchSym = 3;
struct SysDiag {
  long long base, stride;
  int k;                    // 0: the system runs the CSR product
  int off[kBatchSym];       // ascending
};

// Entry of the active list of systems on diagonals: one row block, or two
// consecutive blocks of one system (bl1 >= 0: rows r0..rm-1 are block bl0,
// rm..r1-1 block bl1; otherwise rm = r1), the system, its rows and its
// SysDiag, in one 64-byte load, so the row loads of a workgroup follow from
// that one load.
struct alignas(16) DiaEntry {
  int bl0, bl1, r0, rm;
  int r1, s, sr0, sr1;
  int k, off[kBatchSym];
  long long base, stride;
};
static_assert(sizeof(DiaEntry) == 64, "one 64-byte load");

struct BatchArgs {
  int nsys, nblk, nact;
  SpmvShape m;              // concatenated CSR; m.blk = all row blocks
  const int* bsys;          // system of each block
  const int* sblk;          // nsys+1: first block of each system
  const int* srow;          // nsys+1: first row of each system
  const int4v* act;         // active blocks: {block, first row, end row, system}
  const DiaEntry* act_dia;  // the active blocks whose system takes its products from
  const int4v* act_csr;     // its diagonals (SysDiag), and the others
  int nact_dia, nact_csr;
  const double* val;
  const double* b;
  double* x;
  double* r;
  double* p;
  double* q;
  double* ad;
  double* part;             // [BP_ROWS][nblk]
  State* st;                // one per system
  unsigned long long* clk;  // [0] start; [1 + s] stop of system s (device clock)
  const SysDiag* sdg;       // per system; null: every system runs the CSR product
  double* sd;               // the diagonals of the systems with k > 0
  int* bad;                 // per system: 1 when this solve's val is not symmetric
  Bt32 f;                   // FP32 form
};

// Launch choices of a batched solve (chosen by the batch autotuner).
struct BatchCfg {
  int chunk;      // iterations per replayed graph (FP64)
  int diag_mask;  // bit c: systems of class c on diagonals use the diagonals
  int rows;       // rows per thread of the vector kernels (1 or 2 blocks per workgroup)
};


__device__ __forceinline__ double bmeasure(const State* st, double nrms, double xmin,
                                           double xmax, double rr) {
  if (st->criterion == TFG_CRIT_RELRES) return sqrt(rr) / st->bnorm;
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(xmax - xmin, kTiny);
  return sqrt(nrms) / range;
}

// An active-list entry: {block, first row, end row, system}, one 16-byte
// load, so a workgroup's row loads and its system's state follow from it.
struct ActEntry {
  int bl, r0, r1, s;
};
// Takes the values here (an empty asm), so the loads that produced them are
// issued before any branch that follows: a workgroup's row loads and its
// system's state are then in flight together instead of one after another.
__device__ __forceinline__ void hold1(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void hold1(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void hold1(int v) { asm volatile("" ::"v"(v)); }
template <class... T>
__device__ __forceinline__ void hold(T... v) {
  (hold1(v), ...);
}

// Entry of this workgroup in a list of nact entries, dealt to the XCDs in
// contiguous ranges (workgroup g runs on XCD g % 8); false past the list.
__device__ __forceinline__ bool list_entry(const int4v* list, int nact, ActEntry& e) {
  const int g = blockIdx.x, x = g % kXcds, l = g / kXcds;
  const int first = (int)(((long long)nact * x) / kXcds);
  const int end = (int)(((long long)nact * (x + 1)) / kXcds);
  if (first + l >= end) return false;
  const int4v v = list[first + l];
  e = ActEntry{v.x, v.y, v.z, v.w};
  return true;
}
__device__ __forceinline__ bool active_entry(const BatchArgs& a, ActEntry& e) {
  return list_entry(a.act, a.nact, e);
}

// The vector kernels (p update, x/r update) take NE consecutive entries of
// the XCD's range per workgroup, each thread one row of each (NE = 1 or 2,
// the batch autotuner's rows per thread).  Returns how many exist (0:
// none); a missing one repeats the last, so its loads stay in bounds, and
// stores nothing.
template <int NE>
__device__ __forceinline__ int active_entries(const BatchArgs& a, ActEntry (&e)[NE]) {
  const int g = blockIdx.x, x = g % kXcds, l = g / kXcds;
  const int first = (int)(((long long)a.nact * x) / kXcds);
  const int end = (int)(((long long)a.nact * (x + 1)) / kXcds);
  const int j = first + l * NE;
  if (j >= end) return 0;
  const int n = min(NE, end - j);
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int4v v = a.act[j + min(k, n - 1)];
    e[k] = ActEntry{v.x, v.y, v.z, v.w};
  }
  return n;
}

// Publishes the block totals of the live entries: slot k * NV + c of the
// reduction holds value c of entry k.  Each slot has its own butterfly, so
// an entry's totals are those of a one-entry workgroup, bit for bit.
template <int NV, int NE>
__device__ __forceinline__ void entries_publish(const BatchArgs& a, int row0,
                                                const ActEntry (&e)[NE], const bool (&live)[NE],
                                                double (&v)[NE * NV], const int (&ops)[NE * NV],
                                                double* s_tmp, double* s_out) {
  block_reduce<NE * NV>(v, ops, s_tmp, s_out);
  const int t = (int)threadIdx.x, k = t / NV, c = t % NV;
  if (t < NE * NV && live[k]) a.part[(size_t)(row0 + c) * a.nblk + e[k].bl] = s_out[t];
}

// Sum (min, max) of one system's block partials, fixed order; s_out[k].
template <int NV>
__device__ __forceinline__ void sys_gather(const BatchArgs& a, int row0, int s, const int (&ops)[NV],
                                           double* s_tmp, double* s_out) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = red_identity(ops[k]);
  for (int bl = a.sblk[s] + (int)threadIdx.x; bl < a.sblk[s + 1]; bl += kBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      acc[k] = red_apply(ops[k], acc[k], a.part[(size_t)(row0 + k) * a.nblk + bl]);
  }
  block_reduce<NV>(acc, ops, s_tmp, s_out);
}

// One partial per block: totals of the workgroup at part[(row0+k)*nblk + bl].
template <int NV>
__device__ __forceinline__ void block_publish(const BatchArgs& a, int row0, int bl, double (&v)[NV],
                                              const int (&ops)[NV], double* s_tmp, double* s_out) {
  block_reduce<NV>(v, ops, s_tmp, s_out);
  if (threadIdx.x < NV) a.part[(size_t)(row0 + threadIdx.x) * a.nblk + bl] = s_out[threadIdx.x];
}

__device__ __forceinline__ void stamp(const BatchArgs& a, int s) {
  a.clk[1 + s] = wall_clock64();
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_bt_clock(BatchArgs a) {
  if (threadIdx.x == 0) a.clk[0] = wall_clock64();
}

// sum of the diagonal per block                              Cg.f90:62-73
__global__ __launch_bounds__(kBlock) void k_bt_diag(BatchArgs a) {
  __shared__ double s_tmp[4], s_out[1];
  const int bl = blockIdx.x;
  const int r0 = a.m.blk[bl], r1 = a.m.blk[bl + 1];
  double v[1] = {0.0};
  if (r0 + (int)threadIdx.x < r1) v[0] = a.val[a.m.row[r0 + threadIdx.x]];
  const int ops[1] = {OP_SUM};
  block_publish<1>(a, BP_DIAG, bl, v, ops, s_tmp, s_out);
}

// fn = 1 / (sum_a / n) per system (Cg.f90:75-76), in its fixed block order.
// (A block's diagonal sum above runs as a tree; the system total is the same
// fixed tree for every run.)
__global__ __launch_bounds__(kBlock) void k_bt_fn(BatchArgs a) {
  __shared__ double s_tmp[4], s_out[1];
  const int s = blockIdx.x;
  const int ops[1] = {OP_SUM};
  sys_gather<1>(a, BP_DIAG, s, ops, s_tmp, s_out);
  if (threadIdx.x == 0) {
    const double n = (double)(a.srow[s + 1] - a.srow[s]);
    a.st[s].fn = 1.0 / (s_out[0] / n);
  }
}

// r = b fn - (A fn) x, the scaled diagonal and the 7 start sums per block
template <class E, int PREC>
__global__ __launch_bounds__(kBlock) void k_bt_prologue(BatchArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[7 * 4], s_out[7];
  const int bl = blockIdx.x;
  const double fn = a.st[a.bsys[bl]].fn;
  double v[7] = {0.0, 0.0, kHuge, -kHuge, 0.0, 0.0, 0.0};
  E::template one<true>(
      sm, a.m, a.val, fn, a.x, s_tmp, s_out, bl, [&](int i) { return a.b[i] * fn; },
      [&](int i, double ri, double xi, double di) {
        const double bi = a.b[i] * fn;
        const double z = PREC == PREC_JACOBI ? ri * (1.0 / di) : ri;
        a.r[i] = ri;
        a.ad[i] = di;
        v[0] = v[0] + (bi * bi) / (di * di);
        v[1] = v[1] + (ri * ri) / (di * di);
        v[2] = fmin(v[2], xi);
        v[3] = fmax(v[3], xi);
        v[4] = v[4] + bi * bi;
        v[5] = v[5] + ri * ri;
        v[6] = v[6] + z * ri;
      });
  const int ops[7] = {OP_SUM, OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM, OP_SUM};
  block_publish<7>(a, BP_PRO, bl, v, ops, s_tmp, s_out);
}

// early exits of Cg.f90:95-123 and rho_1, per system
__global__ __launch_bounds__(kBlock) void k_bt_start(BatchArgs a) {
  __shared__ double s_tmp[7 * 4], s_out[7];
  const int s = blockIdx.x;
  const int ops[7] = {OP_SUM, OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM, OP_SUM};
  sys_gather<7>(a, BP_PRO, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  State* st = &a.st[s];
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(s_out[3] - s_out[2], kTiny);
  const double bnorm = sqrt(s_out[4]);
  double bn, res0;
  if (st->criterion == TFG_CRIT_RELRES) {
    bn = bnorm > 0.0 ? 1.0 : 0.0;
    res0 = sqrt(s_out[5]) / bnorm;
  } else {
    bn = sqrt(s_out[0]) / range;
    res0 = sqrt(s_out[1]) / range;
  }
  st->bnorm = bnorm;
  st->iter = 0;
  st->niter = 0;
  st->status = TFG_OK;
  if (bn < st->tol) {
    st->res = 0.0;
    st->done = 1;
    stamp(a, s);
  } else if (res0 < st->tol) {
    st->res = res0;
    st->done = 1;
    stamp(a, s);
  } else {
    st->res = res0;
    st->rho[1] = s_out[6];
    st->done = 0;
  }
}

// p = z + beta p (p = z in a system's first iteration)      Cg.f90:145-171
// (Row loads and the system's state are in flight together; a thread past
// the block loads the block's last row and stores nothing.)
template <int PREC, int NE>
__global__ __launch_bounds__(kBlock) void k_bt_pupdate(BatchArgs a) {
  ActEntry e[NE];
  const int n = active_entries<NE>(a, e);
  if (n == 0) return;
  int i[NE], done[NE], iter[NE];
  double r[NE], pv[NE], d[NE], beta[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    i[k] = min(e[k].r0 + (int)threadIdx.x, e[k].r1 - 1);
    r[k] = a.r[i[k]];
    pv[k] = a.p[i[k]];
    d[k] = PREC == PREC_JACOBI ? a.ad[i[k]] : 1.0;
    const State* st = &a.st[e[k].s];
    done[k] = st->done;
    iter[k] = st->iter;
    beta[k] = st->aux1;
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) hold(r[k], pv[k], d[k], done[k], iter[k], beta[k]);
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    if (k >= n || done[k] || e[k].r0 + (int)threadIdx.x >= e[k].r1) continue;
    const double z = PREC == PREC_JACOBI ? r[k] * (1.0 / d[k]) : r[k];
    a.p[i[k]] = iter[k] == 0 ? z : z + beta[k] * pv[k];
  }
}

// Row i of a system on diagonals (SysDiag): a_ii x_i, the lower entries from
// their mirrors for j = k-1 .. 0, then the upper entries for j = 0 .. k-1:
// the CSR order, so the product is the CSR engine's bit for bit (a missing
// entry adds a zero product; the x index of one is clamped into the system).
// dia_load reads the row's values and the x entries they multiply (the
// system's rows are e.sr0 .. e.sr1-1), dia_hold keeps those loads issued
// ahead of the stop test, dia_sum adds them up.
template <int K, class T>
struct DiaRow {
  T d, x0, lo[K], up[K], xl[K], xu[K];
};
template <int K, class T>
__device__ __forceinline__ void dia_load(const T* sd, const T* x, const DiaEntry& e, int i,
                                         DiaRow<K, T>& w) {
  const long long li = i - e.sr0;
  w.d = sd[li];
  w.x0 = x[i];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int o = e.off[j];
    w.up[j] = sd[(j + 1) * e.stride + li];
    w.lo[j] = sd[(j + 1) * e.stride + li - o];
    w.xu[j] = x[min(i + o, e.sr1 - 1)];
    w.xl[j] = x[max(i - o, e.sr0)];
  }
}
template <int K, class T>
__device__ __forceinline__ void dia_hold(const DiaRow<K, T>& w) {
  hold(w.d, w.x0);
#pragma unroll
  for (int j = 0; j < K; ++j) hold(w.lo[j], w.up[j], w.xl[j], w.xu[j]);
}
template <int K>
__device__ __forceinline__ double dia_sum(const DiaRow<K, double>& w, double fn) {
  double acc = 0.0 + (w.d * fn) * w.x0;
#pragma unroll
  for (int j = K - 1; j >= 0; --j) acc = acc + (w.lo[j] * fn) * w.xl[j];
#pragma unroll
  for (int j = 0; j < K; ++j) acc = acc + (w.up[j] * fn) * w.xu[j];
  return acc;
}

// q = (A fn) p and the block's (p, q)                       Cg.f90:176-195
// Blocks of systems on diagonals (act_dia): no LDS staging, so the kernel
// runs at full occupancy.
template <int K, int NE>
__device__ __forceinline__ void bt_spmv_dia_block(const BatchArgs& a, const DiaEntry& e,
                                                  double* s_tmp, double* s_out) {
  // the rows' loads and the state's are in flight together; rows past a
  // block load its last row and are dropped.  Thread t takes row t of the
  // entry's first block and (NE = 2) row t of its second.
  const int lo[2] = {e.r0, e.rm}, hi[2] = {e.rm, e.r1}, bl[2] = {e.bl0, e.bl1};
  int ti[NE], i[NE];
  DiaRow<K, double> w[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    ti[k] = lo[k] + (int)threadIdx.x;
    i[k] = min(ti[k], max(hi[k], e.r0 + 1) - 1);
    dia_load<K, double>(a.sd + e.base, a.p, e, i[k], w[k]);
  }
  const State* st = &a.st[e.s];
  const int done = st->done;
  const double fn = st->fn;
#pragma unroll
  for (int k = 0; k < NE; ++k) dia_hold<K, double>(w[k]);
  hold(done, fn);
  if (done) return;
  double v[NE];
  int ops[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    v[k] = 0.0;
    ops[k] = OP_SUM;
    if (ti[k] < hi[k]) {
      const double qi = dia_sum<K>(w[k], fn);
      a.q[i[k]] = qi;
      v[k] = v[k] + w[k].x0 * qi;
    }
  }
  block_reduce<NE>(v, ops, s_tmp, s_out);
  const int t = (int)threadIdx.x;
  if (t < NE && bl[t] >= 0) a.part[(size_t)BP_SPMV * a.nblk + bl[t]] = s_out[t];
}
template <int NE>
__global__ __launch_bounds__(kBlock) void k_bt_spmv_dia(BatchArgs a) {
  __shared__ double s_tmp[NE * 4], s_out[NE];
  const int gi = blockIdx.x, x = gi % kXcds, l = gi / kXcds;
  const int first = (int)(((long long)a.nact_dia * x) / kXcds);
  const int end = (int)(((long long)a.nact_dia * (x + 1)) / kXcds);
  if (first + l >= end) return;
  const DiaEntry e = a.act_dia[first + l];
  if (e.k == 1) bt_spmv_dia_block<1, NE>(a, e, s_tmp, s_out);
  else if (e.k == 2) bt_spmv_dia_block<2, NE>(a, e, s_tmp, s_out);
  else bt_spmv_dia_block<3, NE>(a, e, s_tmp, s_out);
}

// The other blocks (act_csr): the CSR row engine E
template <class E>
__global__ __launch_bounds__(kBlock) void k_bt_spmv(BatchArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  ActEntry e;
  if (!list_entry(a.act_csr, a.nact_csr, e)) return;
  const State* st = &a.st[e.s];
  if (st->done) return;
  const double fn = st->fn;
  double v[1] = {0.0};
  E::template one<false>(
      sm, a.m, a.val, fn, a.p, s_tmp, s_out, e.bl, [](int) { return 0.0; },
      [&](int i, double qi, double pi, double) {
        a.q[i] = qi;
        v[0] = v[0] + pi * qi;
      });
  const int ops[1] = {OP_SUM};
  block_publish<1>(a, BP_SPMV, e.bl, v, ops, s_tmp, s_out);
}

// The diagonals of every system with k > 0 from val, one thread per row:
// the diagonal and the upper entries to their slots; every lower entry
// a(i, j) compared bit for bit with a(j, i) from row j, a difference
// setting bad[s] (that system then runs the CSR product in this solve).
__global__ __launch_bounds__(kBlock) void k_bt_sympack(BatchArgs a) {
  const int bl = blockIdx.x;
  const int s = a.bsys[bl];
  const SysDiag g = a.sdg[s];
  if (g.k == 0) return;
  const int i = a.m.blk[bl] + (int)threadIdx.x;
  if (i >= a.m.blk[bl + 1]) return;
  const long long li = i - a.srow[s];
  double* sd = a.sd + g.base;
  bool bad = false;
  for (int j = a.m.row[i]; j < a.m.row[i + 1]; ++j) {
    const int c = a.m.col[j];
    const double v = a.val[j];
    if (c == i) {
      sd[li] = v;
    } else if (c > i) {
      for (int t = 0; t < g.k; ++t)
        if (c - i == g.off[t]) sd[(t + 1) * g.stride + li] = v;
    } else {
      bool found = false;
      for (int t = a.m.row[c]; t < a.m.row[c + 1]; ++t)
        if (a.m.col[t] == i) {
          found = true;
          if (__double_as_longlong(a.val[t]) != __double_as_longlong(v)) bad = true;
        }
      if (!found) bad = true;
    }
  }
  if (bad) atomicOr(&a.bad[s], 1);
}

// alpha = rho / (p, q) per system                            Cg.f90:190-198
__global__ __launch_bounds__(kBlock) void k_bt_alpha(BatchArgs a) {
  __shared__ double s_tmp[4], s_out[1];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done) return;
  const int ops[1] = {OP_SUM};
  sys_gather<1>(a, BP_SPMV, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter + 1;
  const double pq = s_out[0];
  if (pq == 0.0 || !isfinite(pq)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
    stamp(a, s);
    return;
  }
  st->alpha = st->rho[k & 1] / pq;
}

// x += alpha p, r -= alpha q and the block's stopping sums  Cg.f90:204-218
template <int PREC, int NE>
__global__ __launch_bounds__(kBlock) void k_bt_update(BatchArgs a) {
  constexpr int NV = 5;
  __shared__ double s_tmp[NE * NV * 4], s_out[NE * NV];
  ActEntry e[NE];
  const int n = active_entries<NE>(a, e);
  if (n == 0) return;
  int i[NE], done[NE];
  bool live[NE];
  double d[NE], x0[NE], p0[NE], r0[NE], q0[NE], alpha[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    i[k] = min(e[k].r0 + (int)threadIdx.x, e[k].r1 - 1);
    d[k] = a.ad[i[k]];
    x0[k] = a.x[i[k]];
    p0[k] = a.p[i[k]];
    r0[k] = a.r[i[k]];
    q0[k] = a.q[i[k]];
    const State* st = &a.st[e[k].s];
    done[k] = st->done;
    alpha[k] = st->alpha;
  }
  bool any = false;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    hold(d[k], x0[k], p0[k], r0[k], q0[k], done[k], alpha[k]);
    live[k] = k < n && !done[k];
    any = any || live[k];
  }
  if (!any) return;
  double v[NE * NV];
  int ops[NE * NV];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    double* w = v + k * NV;
    w[0] = 0.0;
    w[1] = kHuge;
    w[2] = -kHuge;
    w[3] = 0.0;
    w[4] = 0.0;
    ops[k * NV + 0] = OP_SUM;
    ops[k * NV + 1] = OP_MIN;
    ops[k * NV + 2] = OP_MAX;
    ops[k * NV + 3] = OP_SUM;
    ops[k * NV + 4] = OP_SUM;
    if (!live[k] || e[k].r0 + (int)threadIdx.x >= e[k].r1) continue;
    const double xv = x0[k] + alpha[k] * p0[k];
    const double rv = r0[k] - alpha[k] * q0[k];
    a.x[i[k]] = xv;
    a.r[i[k]] = rv;
    const double z = PREC == PREC_JACOBI ? rv * (1.0 / d[k]) : rv;
    w[0] = (rv * rv) / (d[k] * d[k]);
    w[1] = xv;
    w[2] = xv;
    w[3] = z * rv;
    w[4] = rv * rv;
  }
  entries_publish<NV, NE>(a, BP_UPD, e, live, v, ops, s_tmp, s_out);
}

// stopping test of iteration k (Cg.f90:214-222), next rho and beta
__global__ __launch_bounds__(kBlock) void k_bt_check(BatchArgs a) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done) return;
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  sys_gather<5>(a, BP_UPD, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter + 1;
  const double res = bmeasure(st, s_out[0], s_out[1], s_out[2], s_out[4]);
  st->res = res;
  st->iter = k;
  if (!isfinite(res)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
    stamp(a, s);
  } else if (res < st->tol) {
    st->niter = k;
    st->done = 1;
    stamp(a, s);
  } else {
    st->rho[(k + 1) & 1] = s_out[3];
    st->aux1 = s_out[3] / st->rho[k & 1];
  }
}

// ===========================================================================
// FP32 sweeps inside FP64 refinement (TFG_FP32_IR), system by system.  No
// counterpart in the reference.  Each system runs the single-system form of
// tfg_ir.hip: the CG of Cg.f90 in FP32 on the row-excess matrix (row i of
// the product is s_i x_i + sum_{j != i} a_ij (x_j - x_i), s_i = sum_j a_ij
// fn), FP64 alpha, beta and dot products, the constant vector deflated when
// the system's values are symmetric and nearly singular, and every
// kBt32Check iterations a refinement when due (the FP32 residual fell by
// delta since the last restart, its estimate of the stopping measure is
// below tol, `force` iterations passed, or the cap): x += e in FP64, the
// FP64 true residual, the stopping test on it, the deflation's coarse
// correction and a restart.  Block kernels do the vector work; one
// workgroup per system sweeps that system's block partials and sets its
// scalars.  Per iteration: pupdate | product + (p, q) | alpha | update +
// sums | check (beta, mu, refinement due?); per chunk end: fold | true
// residual | test | restart | restart scalars.
// ===========================================================================

// Conversion (all blocks, once per solve): s_i, the FP32 matrix, 1 / a_ii;
// partials of sum s_i, sum a_ii fn and the double-double sum of b_i fn.
__global__ __launch_bounds__(kBlock) void k_bt32_convert(BatchArgs a) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  __shared__ DD s_dd[kBlock / 64 + 1];
  const int bl = blockIdx.x;
  const int s = a.bsys[bl];
  const double fn = a.st[s].fn;
  const SysDiag g = a.sdg ? a.sdg[s] : SysDiag{};
  const bool dia = g.k > 0 && !a.bad[s];
  double v[2] = {0.0, 0.0};
  DD bs{0.0, 0.0};
  const int i = a.m.blk[bl] + (int)threadIdx.x;
  if (i < a.m.blk[bl + 1]) {
    const int rs = a.m.row[i], re = a.m.row[i + 1];
    const long long li = i - a.srow[s];
    CSum cs{0.0, 0.0};
    for (int j = rs; j < re; ++j) {
      const double aij = a.val[j] * fn;
      cs.add(aij);
      if (j == rs) continue;
      a.f.val32[j] = (float)aij;
      if (dia) {
        const int d = a.m.col[j] - i;
#pragma unroll
        for (int t = 0; t < kBatchSym; ++t)
          if (t < g.k && d == g.off[t]) a.f.sd32[g.base + (t + 1) * g.stride + li] = (float)aij;
      }
    }
    const double sx = cs.s + cs.c;
    const double di = a.val[rs] * fn;
    a.f.val32[rs] = (float)sx;
    if (dia) a.f.sd32[g.base + li] = (float)sx;
    a.f.s64[i] = sx;
    a.f.s32[i] = (float)sx;
    a.f.rd32[i] = (float)(1.0 / di);
    v[0] = sx;
    v[1] = di;
    bs = DD{a.b[i] * fn, 0.0};
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  block_publish<2>(a, BP_IRC, bl, v, ops, s_tmp, s_out);
  const DD t = dd_block(bs, s_dd);
  if (threadIdx.x == 0) {
    a.part[(size_t)(BP_IRC + 2) * a.nblk + bl] = t.hi;
    a.part[(size_t)(BP_IRC + 3) * a.nblk + bl] = t.lo;
  }
}

// Values symmetric bit for bit, for the systems on the CSR product (those on
// diagonals were checked by k_bt_sympack): every lower entry a(i, c) is
// compared with a(c, i) from row c; a difference or a missing mirror sets
// f.asym[s].
__global__ __launch_bounds__(kBlock) void k_bt32_symcheck(BatchArgs a) {
  const int bl = blockIdx.x;
  const int s = a.bsys[bl];
  if (a.sdg && a.sdg[s].k > 0) return;
  const int i = a.m.blk[bl] + (int)threadIdx.x;
  if (i >= a.m.blk[bl + 1]) return;
  bool bad = false;
  for (int j = a.m.row[i]; j < a.m.row[i + 1]; ++j) {
    const int c = a.m.col[j];
    if (c >= i) continue;
    bool found = false;
    for (int t = a.m.row[c]; t < a.m.row[c + 1]; ++t)
      if (a.m.col[t] == i) {
        found = true;
        if (__double_as_longlong(a.val[t]) != __double_as_longlong(a.val[j])) bad = true;
      }
    if (!found) bad = true;
  }
  if (bad) atomicOr(&a.f.asym[s], 1);
}

// Per system: the conversion sums and the deflation choice (automatic for a
// system whose values are symmetric bit for bit and whose row sums add to
// less than 1e-2 of the diagonal's sum, as the single-system form chooses).
__global__ __launch_bounds__(kBlock) void k_bt32_sysinit(BatchArgs a) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  __shared__ DD s_dd[kBlock / 64 + 1];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  const int ops[2] = {OP_SUM, OP_SUM};
  sys_gather<2>(a, BP_IRC, s, ops, s_tmp, s_out);
  DD b{0.0, 0.0};
  for (int bl = a.sblk[s] + (int)threadIdx.x; bl < a.sblk[s + 1]; bl += kBlock)
    b = dd_add(b, DD{a.part[(size_t)(BP_IRC + 2) * a.nblk + bl],
                     a.part[(size_t)(BP_IRC + 3) * a.nblk + bl]});
  b = dd_block(b, s_dd);
  if (threadIdx.x != 0) return;
  const double E = s_out[0], sad = s_out[1];
  const SysDiag g = a.sdg ? a.sdg[s] : SysDiag{};
  const bool sym = g.k > 0 ? !a.bad[s] : !a.f.asym[s];
  const bool defl = E > 0.0 && (st->defl_mode == TFG_DEFLATE_ON ||
                                (st->defl_mode == TFG_DEFLATE_AUTO && sym && E < 1e-2 * sad));
  st->dsum[0] = E;
  st->dsum[1] = sad;
  st->dsum[2] = b.hi;
  st->dsum[3] = b.lo;
  st->defl_e = defl ? E : 0.0;
  st->refine = 1;  // the start runs the restart kernels
  st->kref = 0;
  st->nrefine = 0;
}

// x += e + alpha p (the correction with the last iteration's step) when a
// refinement is due
__global__ __launch_bounds__(kBlock) void k_bt32_fold(BatchArgs a) {
  ActEntry e;
  if (!active_entry(a, e)) return;
  const State* st = &a.st[e.s];
  if (st->done || !st->refine) return;
  const int i = e.r0 + (int)threadIdx.x;
  if (i >= e.r1) return;
  const float ap = (float)st->alpha;
  a.x[i] = a.x[i] + (double)(a.f.e32[i] + ap * a.f.p32[i]);
  a.f.e32[i] = 0.0f;
}

// The FP64 true residual r = b fn - (A fn) x of a refinement, its stopping
// sums and (s, x)
template <class E>
__global__ __launch_bounds__(kBlock) void k_bt32_resid(BatchArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[5 * 4], s_out[5];
  ActEntry e;
  if (!active_entry(a, e)) return;
  const int bl = e.bl;
  const State* st = &a.st[e.s];
  if (st->done || !st->refine) return;
  const double fn = st->fn;
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  E::template one<true>(
      sm, a.m, a.val, fn, a.x, s_tmp, s_out, bl, [&](int i) { return a.b[i] * fn; },
      [&](int i, double ri, double xi, double di) {
        a.r[i] = ri;
        v[0] = v[0] + (ri * ri) / (di * di);
        v[1] = fmin(v[1], xi);
        v[2] = fmax(v[2], xi);
        v[3] = v[3] + ri * ri;
        v[4] = v[4] + a.f.s64[i] * xi;
      });
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  block_publish<5>(a, BP_RES, bl, v, ops, s_tmp, s_out);
}

// Per system: the stopping test on the true residual (niter = the FP32
// iterations so far); otherwise the coarse correction c = w'(b - A x) / E
// = (sum b - (s, x)) / E of the restart
__global__ __launch_bounds__(kBlock) void k_bt32_test(BatchArgs a) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done || !st->refine) return;
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  sys_gather<5>(a, BP_RES, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter;
  const double xmin = s_out[1], xmax = s_out[2], sx = s_out[4];
  const double res = bmeasure(st, s_out[0], xmin, xmax, s_out[3]);
  const int counted = k > 0 ? 1 : 0;  // the start's residual is not a refinement
  st->nrefine = st->nrefine + counted;
  st->res = res;
  if (!isfinite(res) || res < st->tol || k >= st->miter) {
    st->niter = k;
    st->status = isfinite(res) ? TFG_OK : TFG_EBREAKDOWN;
    st->done = 1;
    stamp(a, s);
    return;
  }
  st->kref = k;
  st->xmin = xmin;
  st->xmax = xmax;
  st->corr = st->defl_e > 0.0 ? ((st->dsum[2] - sx) + st->dsum[3]) / st->defl_e : 0.0;
}

// Restart: x += c w, r32 = fl32(r - c s), e = 0, and the sums of the
// restarted residual
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bt32_restart(BatchArgs a) {
  __shared__ double s_tmp[4 * 4], s_out[4];
  ActEntry e;
  if (!active_entry(a, e)) return;
  const int bl = e.bl;
  const State* st = &a.st[e.s];
  if (st->done || !st->refine) return;
  const double c = st->corr;
  const bool defl = st->defl_e > 0.0;
  const bool nrm = st->criterion != TFG_CRIT_RELRES;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const int i = e.r0 + (int)threadIdx.x;
  if (i < e.r1) {
    const float sv = a.f.s32[i], d = a.f.rd32[i];
    if (defl) a.x[i] = a.x[i] + c;
    const float r = (float)(a.r[i] - c * a.f.s64[i]);
    a.f.r32[i] = r;
    a.f.e32[i] = 0.0f;
    const float z = PREC == PREC_JACOBI ? r * d : r;
    v[0] = (double)z * (double)r;
    v[1] = defl ? (double)sv * (double)z : 0.0;
    v[2] = (double)r * (double)r;
    if (nrm) {
      const double t = (double)(r * d);
      v[3] = t * t;
    }
  }
  const int ops[4] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  block_publish<4>(a, BP_UPD, bl, v, ops, s_tmp, s_out);
}

// Per system: the scalars the restarted iteration starts from
__global__ __launch_bounds__(kBlock) void k_bt32_rstart(BatchArgs a) {
  __shared__ double s_tmp[4 * 4], s_out[4];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done || !st->refine) return;
  const int ops[4] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  sys_gather<4>(a, BP_UPD, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter;
  const double rho = s_out[0];
  if (!isfinite(rho)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
    stamp(a, s);
    return;
  }
  st->rho[(k + 1) & 1] = rho;
  st->rref = s_out[2];
  st->aux1 = 0.0;                                                 // beta
  st->aux2 = st->defl_e > 0.0 ? s_out[1] / st->defl_e : 0.0;      // mu
  st->refine = 0;
}

// p = z + beta p - mu (a fresh start after a restart: p = z - mu), with the
// deferred e += alpha_{k-1} p_{k-1} on the p read anyway
template <int PREC, int NE>
__global__ __launch_bounds__(kBlock) void k_bt32_pupdate(BatchArgs a) {
  ActEntry e[NE];
  const int n = active_entries<NE>(a, e);
  if (n == 0) return;
  int i[NE], done[NE], iter[NE], kref[NE];
  float r[NE], pv[NE], ev[NE], d[NE];
  double mu64[NE], beta64[NE], alpha64[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    i[k] = min(e[k].r0 + (int)threadIdx.x, e[k].r1 - 1);
    r[k] = a.f.r32[i[k]];
    pv[k] = a.f.p32[i[k]];
    ev[k] = a.f.e32[i[k]];
    d[k] = PREC == PREC_JACOBI ? a.f.rd32[i[k]] : 1.0f;
    const State* st = &a.st[e[k].s];
    done[k] = st->done;
    iter[k] = st->iter;
    kref[k] = st->kref;
    mu64[k] = st->aux2;
    beta64[k] = st->aux1;
    alpha64[k] = st->alpha;
  }
#pragma unroll
  for (int k = 0; k < NE; ++k)
    hold(r[k], pv[k], ev[k], d[k], done[k], iter[k], kref[k], mu64[k], beta64[k], alpha64[k]);
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    if (k >= n || done[k] || e[k].r0 + (int)threadIdx.x >= e[k].r1) continue;
    const float mu = (float)mu64[k];
    const float z = PREC == PREC_JACOBI ? r[k] * d[k] : r[k];
    if (iter[k] == kref[k]) {  // a fresh start after a restart
      a.f.p32[i[k]] = z - mu;
    } else {
      const float beta = (float)beta64[k], ap = (float)alpha64[k];
      a.f.e32[i[k]] = ev[k] + ap * pv[k];
      a.f.p32[i[k]] = (z + beta * pv[k]) - mu;
    }
  }
}

// Row i of a system on diagonals in the row-excess form (loads by dia_load
// on the FP32 diagonals): s_i x_i, the lower entries (x_{i-d} - x_i) for
// j = k-1 .. 0, then the upper ones: the CSR order of the row-excess CSR
// product, so the same bits (a missing entry is a zero times a finite
// difference)
template <int K>
__device__ __forceinline__ float dia_sum32(const DiaRow<K, float>& w) {
  float acc = w.d * w.x0;
#pragma unroll
  for (int j = K - 1; j >= 0; --j) acc = acc + w.lo[j] * (w.xl[j] - w.x0);
#pragma unroll
  for (int j = 0; j < K; ++j) acc = acc + w.up[j] * (w.xu[j] - w.x0);
  return acc;
}

// q = A32 p and the block's (p, q): blocks of systems on diagonals
template <int K, int NE>
__device__ __forceinline__ void bt32_spmv_dia_block(const BatchArgs& a, const DiaEntry& e,
                                                    double* s_tmp, double* s_out) {
  const int lo[2] = {e.r0, e.rm}, hi[2] = {e.rm, e.r1}, bl[2] = {e.bl0, e.bl1};
  int ti[NE], i[NE];
  DiaRow<K, float> w[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {  // as in bt_spmv_dia_block
    ti[k] = lo[k] + (int)threadIdx.x;
    i[k] = min(ti[k], max(hi[k], e.r0 + 1) - 1);
    dia_load<K, float>(a.f.sd32 + e.base, a.f.p32, e, i[k], w[k]);
  }
  const int done = a.st[e.s].done;
#pragma unroll
  for (int k = 0; k < NE; ++k) dia_hold<K, float>(w[k]);
  hold(done);
  if (done) return;
  double v[NE];
  int ops[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    v[k] = 0.0;
    ops[k] = OP_SUM;
    if (ti[k] < hi[k]) {
      const float qi = dia_sum32<K>(w[k]);
      a.f.q32[i[k]] = qi;
      v[k] = (double)w[k].x0 * (double)qi;
    }
  }
  block_reduce<NE>(v, ops, s_tmp, s_out);
  const int t = (int)threadIdx.x;
  if (t < NE && bl[t] >= 0) a.part[(size_t)BP_SPMV * a.nblk + bl[t]] = s_out[t];
}
template <int NE>
__global__ __launch_bounds__(kBlock) void k_bt32_spmv_dia(BatchArgs a) {
  __shared__ double s_tmp[NE * 4], s_out[NE];
  const int gi = blockIdx.x, x = gi % kXcds, l = gi / kXcds;
  const int first = (int)(((long long)a.nact_dia * x) / kXcds);
  const int end = (int)(((long long)a.nact_dia * (x + 1)) / kXcds);
  if (first + l >= end) return;
  const DiaEntry e = a.act_dia[first + l];
  if (e.k == 1) bt32_spmv_dia_block<1, NE>(a, e, s_tmp, s_out);
  else if (e.k == 2) bt32_spmv_dia_block<2, NE>(a, e, s_tmp, s_out);
  else bt32_spmv_dia_block<3, NE>(a, e, s_tmp, s_out);
}

// the other blocks: the CSR row engine E on the FP32 row-excess values
template <class E>
__global__ __launch_bounds__(kBlock) void k_bt32_spmv(BatchArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  ActEntry e;
  if (!list_entry(a.act_csr, a.nact_csr, e)) return;
  const State* st = &a.st[e.s];
  if (st->done) return;
  double v[1] = {0.0};
  E::template one<false, true>(
      sm, a.m, a.f.val32, 1.0f, a.f.p32, s_tmp, s_out, e.bl, [](int) { return 0.0f; },
      [&](int i, float qi, float pi, float) {
        a.f.q32[i] = qi;
        v[0] = v[0] + (double)pi * (double)qi;
      });
  const int ops[1] = {OP_SUM};
  block_publish<1>(a, BP_SPMV, e.bl, v, ops, s_tmp, s_out);
}

// alpha = rho_k / (p, q) per system (FP64)
__global__ __launch_bounds__(kBlock) void k_bt32_alpha(BatchArgs a) {
  __shared__ double s_tmp[4], s_out[1];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done) return;
  const int ops[1] = {OP_SUM};
  sys_gather<1>(a, BP_SPMV, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter + 1;
  const double pq = s_out[0];
  if (pq == 0.0 || !isfinite(pq)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
    stamp(a, s);
    return;
  }
  st->alpha = st->rho[k & 1] / pq;
}

// r -= alpha q and the sums (z, r), (s, z), (r, r), NRMS of the FP32 residual
template <int PREC, int NE>
__global__ __launch_bounds__(kBlock) void k_bt32_update(BatchArgs a) {
  constexpr int NV = 4;
  __shared__ double s_tmp[NE * NV * 4], s_out[NE * NV];
  ActEntry e[NE];
  const int n = active_entries<NE>(a, e);
  if (n == 0) return;
  int i[NE], crit[NE], done[NE];
  bool live[NE];
  float d[NE], r0[NE], q0[NE], sv[NE];
  double alpha64[NE], defl_e[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    i[k] = min(e[k].r0 + (int)threadIdx.x, e[k].r1 - 1);
    d[k] = a.f.rd32[i[k]];
    r0[k] = a.f.r32[i[k]];
    q0[k] = a.f.q32[i[k]];
    sv[k] = a.f.s32[i[k]];
    const State* st = &a.st[e[k].s];
    done[k] = st->done;
    crit[k] = st->criterion;
    alpha64[k] = st->alpha;
    defl_e[k] = st->defl_e;
  }
  bool any = false;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    hold(d[k], r0[k], q0[k], sv[k], done[k], crit[k], alpha64[k], defl_e[k]);
    live[k] = k < n && !done[k];
    any = any || live[k];
  }
  if (!any) return;
  double v[NE * NV];
  int ops[NE * NV];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    double* w = v + k * NV;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      w[c] = 0.0;
      ops[k * NV + c] = OP_SUM;
    }
    if (!live[k] || e[k].r0 + (int)threadIdx.x >= e[k].r1) continue;
    const float a32 = (float)alpha64[k];
    const float r = r0[k] - a32 * q0[k];
    a.f.r32[i[k]] = r;
    const float z = PREC == PREC_JACOBI ? r * d[k] : r;
    w[0] = (double)z * (double)r;
    w[1] = defl_e[k] > 0.0 ? (double)sv[k] * (double)z : 0.0;
    w[2] = (double)r * (double)r;
    if (crit[k] != TFG_CRIT_RELRES) {
      const double t = (double)(r * d[k]);
      w[3] = t * t;
    }
  }
  entries_publish<NV, NE>(a, BP_UPD, e, live, v, ops, s_tmp, s_out);
}

// Per system, end of iteration k: beta and mu of the next iteration and, at
// a chunk end (every kBt32Check iterations) or the cap, whether a
// refinement is due
__global__ __launch_bounds__(kBlock) void k_bt32_check(BatchArgs a) {
  __shared__ double s_tmp[4 * 4], s_out[4];
  const int s = blockIdx.x;
  State* st = &a.st[s];
  if (st->done) return;
  const int ops[4] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  sys_gather<4>(a, BP_UPD, s, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const int k = st->iter + 1;
  st->iter = k;
  const double rho = s_out[0], rr = s_out[2];
  if (!isfinite(rho)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
    stamp(a, s);
    return;
  }
  st->aux1 = rho / st->rho[k & 1];
  st->rho[(k + 1) & 1] = rho;
  st->aux2 = st->defl_e > 0.0 ? s_out[1] / st->defl_e : 0.0;
  if (k % kBt32Check == 0 || k >= st->miter) {
    double est;
    if (st->criterion == TFG_CRIT_RELRES) {
      est = sqrt(rr) / st->bnorm;
    } else {
      const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(st->xmax - st->xmin, kTiny);
      est = sqrt(s_out[3]) / range;
    }
    st->refine = (rr <= st->delta * st->delta * st->rref || est < st->tol ||
                  (st->force > 0 && k - st->kref >= st->force) || k >= st->miter ||
                  !isfinite(rr)) ? 1 : 0;
  }
}

// x += e of every system that broke down (no step pending); all blocks
__global__ __launch_bounds__(kBlock) void k_bt32_fold_e(BatchArgs a) {
  const int bl = blockIdx.x;
  const State* st = &a.st[a.bsys[bl]];
  if (st->status != TFG_EBREAKDOWN) return;
  const int i = a.m.blk[bl] + threadIdx.x;
  if (i >= a.m.blk[bl + 1]) return;
  a.x[i] = a.x[i] + (double)a.f.e32[i];
  a.f.e32[i] = 0.0f;
}

}  // namespace tfg

// ===========================================================================
// Host side
// ===========================================================================

struct tfg_batch {
  int device = 0, nsys = 0, nt = 0, nnz = 0, nblk = 0, cu_count = 0;
  hipStream_t stream = nullptr;
  std::vector<int> h_srow, h_sblk;  // host copies of the system offsets
  std::vector<int> h_blk, h_bsys;   // host copies of the row blocks and their systems
  int *row = nullptr, *col = nullptr, *blk = nullptr, *bsys = nullptr, *sblk = nullptr,
      *srow = nullptr;
  tfg::int4v* act = nullptr;        // active list entries {block, first row, end row, system}
  double *val = nullptr, *b = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr,
         *ad = nullptr, *part = nullptr;
  tfg::State* st = nullptr;
  tfg::State* hst = nullptr;          // pinned, 2 x nsys (polling, double-buffered)
  tfg::SysDiag* sdg = nullptr;        // per system (device); null: no system on diagonals
  double* sd = nullptr;               // their diagonals (zeroed at the first solve)
  long long sd_len = 0;
  bool sd_zeroed = false;
  int* bad = nullptr;                 // per system, per solve
  unsigned long long* clk = nullptr;  // device, nsys + 1
  std::vector<double> done_ms;        // stop time of each system, last solve
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_poll[2] = {nullptr, nullptr};
  tfg::Bt32 f{};                      // FP32 form (allocated on first use)
  bool sd32_zeroed = false;
  std::vector<uint8_t> h_scls;        // size class of each system (0: < 32768 rows, 1: < 262144, 2)
  std::vector<int> h_sdk;             // diagonals of each system (SysDiag::k; 0: CSR rows only)
  std::vector<int> h_bad;             // bad flags of the last solve, read back from the device
  tfg::DiaEntry* act_dia = nullptr;   // the active list split by product
  tfg::int4v* act_csr = nullptr;
  std::vector<tfg::SysDiag> h_sdg;    // host copy of sdg
  bool class_diag[3] = {false, false, false};  // a class holds a system on diagonals
  // batch autotuner, per precision (FP64, FP32): choice, probes, best probe
  bool tuned[2] = {false, false};
  tfg::BatchCfg cfg[2] = {{16, 7, 1}, {16, 7, 1}};
  int probes[2] = {0, 0};
  double probe_ms[2] = {0.0, 0.0};
  double* x0 = nullptr;               // the caller's x0 while probing
  std::vector<unsigned long long> clk_host;  // device clock stamps of the last solve
};

namespace {

using tfg::fail;

// Whether system s (1-based CSR of n rows) lies on its diagonal and at most
// kBatchSym upper diagonals and their mirrors, each row stored diagonal
// first and then by ascending column; g.k = 0 if not.
void sys_diag(int n, const int* row, const int* col, tfg::SysDiag& g) {
  g = tfg::SysDiag{};
  std::vector<int> off;
  long long nup = 0, nlow = 0;
  for (int i = 0; i < n; ++i) {
    const int rs = row[i] - 1, re = row[i + 1] - 1;
    if (re <= rs || col[rs] - 1 != i) return;
    for (int j = rs + 1; j < re; ++j) {
      if (j > rs + 1 && !(col[j - 1] < col[j])) return;
      const int d = col[j] - 1 - i;
      if (d > 0) {
        ++nup;
        if (std::find(off.begin(), off.end(), d) == off.end()) {
          if ((int)off.size() == tfg::kBatchSym) return;
          off.push_back(d);
        }
        // the mirror (i + d, i) in row i + d
        const int c = i + d;
        bool found = false;
        for (int t = row[c] - 1; t < row[c + 1] - 1 && !found; ++t) found = col[t] - 1 == i;
        if (!found) return;
      } else {
        ++nlow;
      }
    }
  }
  if (off.empty() || nup != nlow) return;
  std::sort(off.begin(), off.end());
  g.k = (int)off.size();
  for (int j = 0; j < g.k; ++j) g.off[j] = off[j];
  const long long gap = ((long long)off.back() + 2 + 1) & ~1LL;
  g.stride = ((long long)n + gap + 1) & ~1LL;
}

void batch_free(tfg_batch* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* dev[] = {h->row, h->col, h->blk, h->bsys, h->sblk, h->srow, h->act, h->val, h->b,
                 h->x, h->r, h->p, h->q, h->ad, h->part, h->st, h->clk, h->sdg, h->sd, h->bad,
                 h->f.val32, h->f.sd32, h->f.s64, h->f.s32, h->f.rd32, h->f.r32, h->f.p32,
                 h->f.q32, h->f.e32, h->f.asym, h->act_dia, h->act_csr, h->x0};
  for (void* d : dev)
    if (d) (void)hipFree(d);
  if (h->hst) (void)hipHostFree(h->hst);
  for (hipEvent_t e : {h->ev0, h->ev1, h->ev_poll[0], h->ev_poll[1]})
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  (void)hipGetLastError();  // release errors are dropped, not left for later calls
  delete h;
}

}  // namespace

namespace {

using tfg::BatchCfg;

// One batched solve from h->x with the launch choices `cfg` (the graph's
// iterations and the size classes whose systems on diagonals take their
// products from the diagonals); leaves the final states in h->hst[0..nsys),
// the device clock stamps in h->clk_host and the device time in *ms.
int batch_run(tfg_batch* h, int pr, bool fp32, const tfg_options& o, const double* dval,
              const double* db, int miter, double tol, const double* norm, const BatchCfg& cfg,
              float* ms) {
  using namespace tfg;
  // per-system state
  for (int s = 0; s < h->nsys; ++s) {
    State& q = h->hst[s];
    memset(&q, 0, sizeof(State));
    q.tol = tol;
    q.criterion = o.criterion;
    q.has_norm = norm ? 1 : 0;
    q.norm = norm ? norm[s] : 0.0;
    q.miter = std::max(miter, 0);
    q.delta = o.inner_tol > 0.0 ? o.inner_tol : 1e-3;
    q.force = std::max(o.inner_max, 0);
    q.defl_mode = o.deflation;
  }
  HIP_TRY(hipMemcpyAsync(h->st, h->hst, sizeof(State) * (size_t)h->nsys, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));

  BatchArgs a{};
  a.nsys = h->nsys;
  a.nblk = h->nblk;
  a.m.n = h->nt;
  a.m.row = h->row;
  a.m.col = h->col;
  a.m.blk = h->blk;
  a.m.nblk = h->nblk;
  a.bsys = h->bsys;
  a.sblk = h->sblk;
  a.srow = h->srow;
  a.act = h->act;
  a.val = dval;
  a.b = db;
  a.x = h->x;
  a.r = h->r;
  a.p = h->p;
  a.q = h->q;
  a.ad = h->ad;
  a.part = h->part;
  a.st = h->st;
  a.clk = h->clk;
  // Systems on diagonals take their products from the diagonals, refilled
  // from val here; TFG_BATCH_SPMV=csr keeps every system on the CSR product
  // (the same bits).
  const char* force = getenv("TFG_BATCH_SPMV");
  if (h->sdg && !(force && std::string(force) == "csr")) {
    if (!h->sd_zeroed) {  // entries a row lacks and the gaps stay zero
      HIP_TRY(hipMemsetAsync(h->sd, 0, sizeof(double) * (size_t)h->sd_len, h->stream));
      h->sd_zeroed = true;
    }
    if (fp32 && !h->sd32_zeroed) {
      HIP_TRY(hipMemsetAsync(h->f.sd32, 0, sizeof(float) * (size_t)h->sd_len, h->stream));
      h->sd32_zeroed = true;
    }
    HIP_TRY(hipMemsetAsync(h->bad, 0, sizeof(int) * (size_t)h->nsys, h->stream));
    a.sdg = h->sdg;
    a.sd = h->sd;
    a.bad = h->bad;
  }
  a.f = h->f;
  const bool vec = ((uintptr_t)dval % 16) == 0 && h->nnz >= 4;
  const bool vec32 = h->nnz >= 4;  // val32 is the library's, 16-byte aligned
  const hipStream_t S = h->stream;
  const int B = kBlock;

#define TFG_BT_ENGINE(CALL)                                                 \
  do {                                                                      \
    if (vec) { using E = OneBlock<kCap, double>; CALL; }               \
    else { using E = ScalarStage<kCap, double>; CALL; }                     \
  } while (0)

  HIP_TRY(hipEventRecord(h->ev0, S));
  k_bt_clock<<<1, B, 0, S>>>(a);
  if (a.sdg) k_bt_sympack<<<h->nblk, B, 0, S>>>(a);
  k_bt_diag<<<h->nblk, B, 0, S>>>(a);
  k_bt_fn<<<h->nsys, B, 0, S>>>(a);
  if (pr == PREC_JACOBI) TFG_BT_ENGINE((k_bt_prologue<E, PREC_JACOBI><<<h->nblk, B, 0, S>>>(a)));
  else TFG_BT_ENGINE((k_bt_prologue<E, PREC_NONE><<<h->nblk, B, 0, S>>>(a)));
  k_bt_start<<<h->nsys, B, 0, S>>>(a);
  if (fp32) {
    k_bt32_symcheck<<<h->nblk, B, 0, S>>>(a);
    k_bt32_convert<<<h->nblk, B, 0, S>>>(a);
    k_bt32_sysinit<<<h->nsys, B, 0, S>>>(a);
  }
  HIP_TRY(hipGetLastError());

  // Which systems take their products from their diagonals: on diagonals,
  // values symmetric in this solve (the flags k_bt_sympack set), class
  // enabled by cfg.diag_mask.
  std::vector<char> dia(h->nsys, 0);
  if (a.sdg) {
    h->h_bad.resize(h->nsys);
    HIP_TRY(hipMemcpyAsync(h->h_bad.data(), h->bad, sizeof(int) * (size_t)h->nsys,
                           hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    for (int s = 0; s < h->nsys; ++s)
      dia[s] = h->h_sdk[s] > 0 && !h->h_bad[s] && ((cfg.diag_mask >> h->h_scls[s]) & 1);
  }
  // Active lists (every block of the running systems; all of them, those on
  // diagonals, the others), rebuilt from the polled state as systems stop.
  std::vector<int4v> act, act_c;
  std::vector<DiaEntry> act_d;
  // Rows per thread of the block kernels (the autotuner's choice, or
  // TFG_BATCH_ROWS): two takes two blocks per workgroup in the vector
  // kernels and two consecutive blocks of a system per diagonal entry.
  const char* force_rows = getenv("TFG_BATCH_ROWS");
  const bool two = force_rows && (*force_rows == '1' || *force_rows == '2') ? *force_rows == '2'
                                                                            : cfg.rows == 2;
  auto build_lists = [&](const State* ps) {
    act.clear();
    act_d.clear();
    act_c.clear();
    for (int s = 0; s < h->nsys; ++s) {
      if (ps && ps[s].done) continue;
      const int b0 = h->h_sblk[s], b1 = h->h_sblk[s + 1];
      for (int t = b0; t < b1; ++t) {
        int4v v;
        v.x = t;
        v.y = h->h_blk[t];
        v.z = h->h_blk[t + 1];
        v.w = s;
        act.push_back(v);
        if (!dia[s]) act_c.push_back(v);
      }
      if (!dia[s]) continue;
      const SysDiag& g = h->h_sdg[s];
      for (int t = b0; t < b1; t += two ? 2 : 1) {
        const bool pair = two && t + 1 < b1;
        DiaEntry d{};
        d.bl0 = t;
        d.bl1 = pair ? t + 1 : -1;
        d.r0 = h->h_blk[t];
        d.rm = h->h_blk[t + 1];
        d.r1 = pair ? h->h_blk[t + 2] : d.rm;
        d.s = s;
        d.sr0 = h->h_srow[s];
        d.sr1 = h->h_srow[s + 1];
        d.k = g.k;
        for (int j = 0; j < kBatchSym; ++j) d.off[j] = g.off[j];
        d.base = g.base;
        d.stride = g.stride;
        act_d.push_back(d);
      }
    }
  };
  auto upload_lists = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(h->act, act.data(), sizeof(int4v) * act.size(), hipMemcpyHostToDevice, S));
    if (!act_d.empty())
      HIP_TRY(hipMemcpyAsync(h->act_dia, act_d.data(), sizeof(DiaEntry) * act_d.size(),
                             hipMemcpyHostToDevice, S));
    if (!act_c.empty())
      HIP_TRY(hipMemcpyAsync(h->act_csr, act_c.data(), sizeof(int4v) * act_c.size(),
                             hipMemcpyHostToDevice, S));
    HIP_TRY(hipStreamSynchronize(S));
    a.nact = (int)act.size();
    a.nact_dia = (int)act_d.size();
    a.nact_csr = (int)act_c.size();
    return TFG_OK;
  };
  a.act_dia = h->act_dia;
  a.act_csr = h->act_csr;
  build_lists(nullptr);
  { const int ru = upload_lists(); if (ru != TFG_OK) return ru; }

#define TFG_BT32_ENGINE(CALL)                                               \
  do {                                                                      \
    if (vec32) { using E = OneBlock<kCap, float>; CALL; }                   \
    else { using E = ScalarStage<kCap, float>; CALL; }                      \
  } while (0)

  const bool use_graph = o.path != TFG_PATH_KERNELS;
  const int chunk = fp32 ? kBt32Check : cfg.chunk;
  hipGraphExec_t exec = nullptr;
  // FP32 form: the true residual, test and restart of the systems whose
  // refinement is due (all of them at the start)
  auto enqueue_refine = [&](int G) {
    TFG_BT_ENGINE((k_bt32_resid<E><<<G, B, 0, S>>>(a)));
    k_bt32_test<<<a.nsys, B, 0, S>>>(a);
    if (pr == PREC_JACOBI) k_bt32_restart<PREC_JACOBI><<<G, B, 0, S>>>(a);
    else k_bt32_restart<PREC_NONE><<<G, B, 0, S>>>(a);
    k_bt32_rstart<<<a.nsys, B, 0, S>>>(a);
  };
  // the vector kernels with cfg.rows entries per workgroup, grid Gv
#define TFG_BT_VEC(KERNEL)                                                            \
  do {                                                                                \
    if (two && pr == PREC_JACOBI) KERNEL<PREC_JACOBI, 2><<<Gv, B, 0, S>>>(a);          \
    else if (two) KERNEL<PREC_NONE, 2><<<Gv, B, 0, S>>>(a);                            \
    else if (pr == PREC_JACOBI) KERNEL<PREC_JACOBI, 1><<<Gv, B, 0, S>>>(a);            \
    else KERNEL<PREC_NONE, 1><<<Gv, B, 0, S>>>(a);                                     \
  } while (0)
  auto enqueue_chunk = [&](int count) -> hipError_t {
    const int per = (a.nact + kXcds - 1) / kXcds;
    const int G = kXcds * per, Gv = kXcds * (two ? (per + 1) / 2 : per);
    const int Gd = kXcds * ((a.nact_dia + kXcds - 1) / kXcds);
    const int Gc = kXcds * ((a.nact_csr + kXcds - 1) / kXcds);
    if (fp32) {
      for (int c = 0; c < count; ++c) {
        TFG_BT_VEC(k_bt32_pupdate);
        if (Gd && two) k_bt32_spmv_dia<2><<<Gd, B, 0, S>>>(a);
        else if (Gd) k_bt32_spmv_dia<1><<<Gd, B, 0, S>>>(a);
        if (Gc) TFG_BT32_ENGINE((k_bt32_spmv<E><<<Gc, B, 0, S>>>(a)));
        k_bt32_alpha<<<a.nsys, B, 0, S>>>(a);
        TFG_BT_VEC(k_bt32_update);
        k_bt32_check<<<a.nsys, B, 0, S>>>(a);
      }
      k_bt32_fold<<<G, B, 0, S>>>(a);
      enqueue_refine(G);
      return hipGetLastError();
    }
    for (int c = 0; c < count; ++c) {
      TFG_BT_VEC(k_bt_pupdate);
      if (Gd && two) k_bt_spmv_dia<2><<<Gd, B, 0, S>>>(a);
      else if (Gd) k_bt_spmv_dia<1><<<Gd, B, 0, S>>>(a);
      if (Gc) TFG_BT_ENGINE((k_bt_spmv<E><<<Gc, B, 0, S>>>(a)));
      k_bt_alpha<<<a.nsys, B, 0, S>>>(a);
      TFG_BT_VEC(k_bt_update);
      k_bt_check<<<a.nsys, B, 0, S>>>(a);
    }
    return hipGetLastError();
  };
  if (fp32) {  // the start: FP64 residual, test, first restart
    enqueue_refine(kXcds * ((a.nact + kXcds - 1) / kXcds));
    HIP_TRY(hipGetLastError());
  }
  auto build_graph = [&]() -> int {
    if (exec) {
      HIP_TRY(hipGraphExecDestroy(exec));
      exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(S, hipStreamCaptureModeThreadLocal));
    const hipError_t e1 = enqueue_chunk(chunk);
    hipError_t e2 = hipStreamEndCapture(S, &g);
    if (e1 != hipSuccess) e2 = e1;
    if (e2 == hipSuccess) e2 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    HIP_TRY(e2);
    return TFG_OK;
  };
  int rc = TFG_OK;
  if (use_graph) rc = build_graph();
  if (rc != TFG_OK) return rc;

  // Chunks of iterations; the state of chunk c is read while chunk c+1
  // runs (stopped systems are masked on the device, so the lag only costs
  // the masked launches of one chunk).
  int k = 0, slot = 0, prev = -1;
  while (k < miter) {
    const int cnt = std::min(chunk, miter - k);
    if (use_graph && cnt == chunk) HIP_TRY(hipGraphLaunch(exec, S));
    else HIP_TRY(enqueue_chunk(cnt));
    k += cnt;
    State* poll = h->hst + (size_t)slot * h->nsys;
    HIP_TRY(hipMemcpyAsync(poll, h->st, sizeof(State) * (size_t)h->nsys, hipMemcpyDeviceToHost, S));
    HIP_TRY(hipEventRecord(h->ev_poll[slot], S));
    if (prev >= 0) {
      HIP_TRY(hipEventSynchronize(h->ev_poll[prev]));
      const State* ps = h->hst + (size_t)prev * h->nsys;
      int running = 0, still = 0;
      for (int s = 0; s < h->nsys; ++s) {
        if (ps[s].done) continue;
        ++running;
        still += h->h_sblk[s + 1] - h->h_sblk[s];
      }
      if (!running) break;
      // compaction once an eighth of the active blocks have stopped (the
      // lists are rebuilt only then; the copies are stream-ordered after
      // the chunk already queued)
      if (still * 8 <= a.nact * 7) {
        build_lists(ps);
        rc = upload_lists();
        if (rc != TFG_OK) return rc;
        if (use_graph) {
          rc = build_graph();
          if (rc != TFG_OK) return rc;
        }
      }
    }
    prev = slot;
    slot ^= 1;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (fp32) {  // the correction of systems that broke down
    k_bt32_fold_e<<<h->nblk, B, 0, S>>>(a);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(h->ev1, S));
  HIP_TRY(hipMemcpyAsync(h->hst, h->st, sizeof(State) * (size_t)h->nsys, hipMemcpyDeviceToHost, S));
  h->clk_host.resize(h->nsys + 1);
  HIP_TRY(hipMemcpyAsync(h->clk_host.data(), h->clk, sizeof(unsigned long long) * h->clk_host.size(),
                         hipMemcpyDeviceToHost, S));
  HIP_TRY(hipStreamSynchronize(S));
  HIP_TRY(hipEventElapsedTime(ms, h->ev0, h->ev1));
  return TFG_OK;
#undef TFG_BT_ENGINE
#undef TFG_BT32_ENGINE
#undef TFG_BT_VEC
}

// The batch autotuner: short timed probes (the first kTuneIters iterations
// from the caller's x0, restored before every probe and before the solve)
// pick the graph's iterations (FP64: 8, 16 or 32; the FP32 form refines
// every 8 and keeps 8) and, per size class present, whether systems on
// diagonals take their products from the diagonals or the CSR rows.  Every
// choice gives the same bits, so a tuned solve is reproducible; the result
// is kept per handle and precision.  TFG_BATCH_TUNE=0 keeps the defaults
// (16 iterations, diagonals for every class).
constexpr int kTuneIters = 32;

int batch_tune(tfg_batch* h, int pr, bool fp32, const tfg_options& o, const double* dval,
               const double* db, int miter, double tol, const double* norm) {
  const int p = fp32 ? 1 : 0;
  if (h->tuned[p]) return TFG_OK;
  h->cfg[p] = BatchCfg{16, 7, 1};
  h->probes[p] = 0;
  h->probe_ms[p] = 0.0;
  const char* env = getenv("TFG_BATCH_TUNE");
  if ((env && *env == '0') || miter < kTuneIters) {
    h->tuned[p] = true;
    return TFG_OK;
  }
  const size_t N = (size_t)h->nt;
  if (!h->x0) HIP_TRY(hipMalloc(&h->x0, sizeof(double) * N));
  HIP_TRY(hipMemcpyAsync(h->x0, h->x, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
  double best = 1e300;
  BatchCfg pick = h->cfg[p];
  auto probe = [&](const BatchCfg& c) -> int {
    HIP_TRY(hipMemcpyAsync(h->x, h->x0, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
    float ms = 0.f;
    const int rc = batch_run(h, pr, fp32, o, dval, db, kTuneIters, tol, norm, c, &ms);
    if (rc != TFG_OK) return rc;
    ++h->probes[p];
    if (ms < best) {
      best = ms;
      pick = c;
    }
    return TFG_OK;
  };
  for (int ch : fp32 ? std::vector<int>{16} : std::vector<int>{8, 16, 32}) {
    const int rc = probe(BatchCfg{ch, 7, 1});
    if (rc != TFG_OK) return rc;
  }
  if (h->sdg) {
    const BatchCfg base = pick;
    for (int c = 0; c < 3; ++c) {
      if (!h->class_diag[c]) continue;  // no system on diagonals in this class
      const int rc = probe(BatchCfg{base.chunk, base.diag_mask & ~(1 << c), 1});
      if (rc != TFG_OK) return rc;
    }
  }
  {  // two rows per thread in the vector kernels
    const BatchCfg base = pick;
    const int rc = probe(BatchCfg{base.chunk, base.diag_mask, 2});
    if (rc != TFG_OK) return rc;
  }
  h->cfg[p] = pick;
  h->probe_ms[p] = best;
  h->tuned[p] = true;
  HIP_TRY(hipMemcpyAsync(h->x, h->x0, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
  return TFG_OK;
}

}  // namespace

extern "C" {

int tfg_batch_create(int nsys, const int* n, const int* nnz, const int* row, const int* col,
                     const int* dia, tfg_batch** handle) {
  if (!handle || !n || !nnz || !row || !col || !dia) return fail(TFG_EINVAL, "null argument");
  *handle = nullptr;
  if (nsys < 1) return fail(TFG_EINVAL, "nsys=%d", nsys);
  long long NT = 0, NNZ = 0;
  for (int s = 0; s < nsys; ++s) {
    if (n[s] < 1 || nnz[s] < n[s]) return fail(TFG_EINVAL, "system %d: n=%d nnz=%d", s, n[s], nnz[s]);
    NT += n[s];
    NNZ += nnz[s];
  }
  if (NT > INT_MAX / 2 || NNZ > INT_MAX - 64) return fail(TFG_EINVAL, "batch too large for int32 indices");
  // validate every system and build the global arrays
  std::vector<int> grow((size_t)NT + 1), gcol((size_t)NNZ), blk, bsys, sblk, srow;
  std::vector<tfg::SysDiag> sdg(nsys);
  long long sd_len = 0;
  int nsd = 0;
  {
    long long ro = 0, co = 0, rpos = 0, dpos = 0;
    std::vector<int> b1;
    blk.push_back(0);
    sblk.push_back(0);
    srow.push_back(0);
    for (int s = 0; s < nsys; ++s) {
      const int* rs = row + rpos;
      const int rc = tfg::check_csr(n[s], nnz[s], rs, col + co, dia + dpos);
      if (rc != TFG_OK) return fail(rc, "system %d: %s", s, tfg_last_error());
      sys_diag(n[s], rs, col + co, sdg[s]);
      if (sdg[s].k > 0) {
        sdg[s].base = sd_len;
        sd_len += (sdg[s].k + 1) * sdg[s].stride;
        ++nsd;
      }
      for (int i = 0; i < n[s]; ++i) grow[ro + i] = (int)(co + rs[i] - 1);
      for (int j = 0; j < nnz[s]; ++j) gcol[co + j] = (int)(ro + col[co + j] - 1);
      tfg::make_blocks(rs, n[s], tfg::kCap, b1);
      for (size_t t = 1; t < b1.size(); ++t) {
        blk.push_back((int)(ro + b1[t]));
        bsys.push_back(s);
      }
      sblk.push_back((int)bsys.size());
      ro += n[s];
      co += nnz[s];
      srow.push_back((int)ro);
      rpos += n[s] + 1;
      dpos += n[s];
    }
    grow[NT] = (int)NNZ;
  }
  tfg_batch* h = new tfg_batch();
  h->nsys = nsys;
  h->nt = (int)NT;
  h->nnz = (int)NNZ;
  h->nblk = (int)bsys.size();
  h->h_srow = srow;
  h->h_sblk = sblk;
  h->h_bsys = bsys;
  h->h_blk = blk;
  h->done_ms.assign(nsys, 0.0);
  int rc = TFG_OK;
  do {
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    const size_t N = (size_t)NT;
    auto dmal = [&](void** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
    dmal((void**)&h->row, sizeof(int) * (N + 1));
    dmal((void**)&h->col, sizeof(int) * (size_t)NNZ);
    dmal((void**)&h->blk, sizeof(int) * blk.size());
    dmal((void**)&h->bsys, sizeof(int) * bsys.size());
    dmal((void**)&h->sblk, sizeof(int) * sblk.size());
    dmal((void**)&h->srow, sizeof(int) * srow.size());
    dmal((void**)&h->act, sizeof(tfg::int4v) * bsys.size());
    for (double** v : {&h->x, &h->r, &h->p, &h->q, &h->ad}) dmal((void**)v, sizeof(double) * N);
    dmal((void**)&h->part, sizeof(double) * (size_t)tfg::BP_ROWS * bsys.size());
    dmal((void**)&h->st, sizeof(tfg::State) * (size_t)nsys);
    dmal((void**)&h->clk, sizeof(unsigned long long) * (size_t)(nsys + 1));
    dmal((void**)&h->act_dia, sizeof(tfg::DiaEntry) * bsys.size());
    dmal((void**)&h->act_csr, sizeof(tfg::int4v) * bsys.size());
    if (nsd > 0) {
      dmal((void**)&h->sdg, sizeof(tfg::SysDiag) * (size_t)nsys);
      dmal((void**)&h->sd, sizeof(double) * (size_t)sd_len);
      dmal((void**)&h->bad, sizeof(int) * (size_t)nsys);
      h->sd_len = sd_len;
    }
    if (e != hipSuccess) { rc = fail(TFG_ENOMEM, "batch allocation: %s", hipGetErrorString(e)); break; }
    e = hipHostMalloc((void**)&h->hst, 2 * sizeof(tfg::State) * (size_t)nsys, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipEventCreate(&h->ev_poll[0]);
    if (e == hipSuccess) e = hipEventCreate(&h->ev_poll[1]);
    auto up = [&](void* d, const void* s, size_t bytes) {
      if (e == hipSuccess) e = hipMemcpy(d, s, bytes, hipMemcpyHostToDevice);
    };
    up(h->row, grow.data(), sizeof(int) * (N + 1));
    up(h->col, gcol.data(), sizeof(int) * (size_t)NNZ);
    up(h->blk, blk.data(), sizeof(int) * blk.size());
    up(h->bsys, bsys.data(), sizeof(int) * bsys.size());
    up(h->sblk, sblk.data(), sizeof(int) * sblk.size());
    up(h->srow, srow.data(), sizeof(int) * srow.size());
    if (h->sdg) up(h->sdg, sdg.data(), sizeof(tfg::SysDiag) * sdg.size());
    {
      // size classes of the autotuner: < 32768 rows, < 262144, larger
      h->h_scls.resize(nsys);
      h->h_sdk.resize(nsys);
      for (int q = 0; q < nsys; ++q) {
        h->h_scls[q] = n[q] < (1 << 15) ? 0 : (n[q] < (1 << 18) ? 1 : 2);
        h->h_sdk[q] = sdg[q].k;
        h->h_sdg.push_back(sdg[q]);
        if (sdg[q].k > 0) h->class_diag[h->h_scls[q]] = true;
      }
    }
    if (e != hipSuccess) { rc = fail(TFG_EHIP, "batch upload: %s", hipGetErrorString(e)); break; }
  } while (0);
  if (rc != TFG_OK) {
    batch_free(h);
    return rc;
  }
  *handle = h;
  return TFG_OK;
}

int tfg_batch_destroy(tfg_batch* h) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  batch_free(h);
  return TFG_OK;
}

int tfg_batch_solve(tfg_batch* h, const char* solver, int solver_len, const char* prec,
                    int prec_len, const double* val, double* x, double* b, int miter, int* niter,
                    double tol, double* fin_res, const double* norm, int device_ptrs,
                    const tfg_options* opt, tfg_report* rep) {
  using namespace tfg;
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (!val || !x || !b || !niter || !fin_res) return fail(TFG_EINVAL, "null argument");
  const std::string sv = fstring(solver, solver_len);
  if (sv != "cg" && sv != "bicg") return fail(TFG_EINVAL, "Unknown native solver: %s", sv.c_str());
  const std::string pv = fstring(prec, prec_len);
  const int pr = pv == "jacobi" ? PREC_JACOBI : (pv == "icc" ? PREC_ICC : PREC_NONE);
  if (pr == PREC_ICC) return fail(TFG_ENOTSUP, "preconditioner 'icc' is not available on the GPU path yet");
  tfg_options o{};
  if (opt) o = *opt;
  if (o.criterion != TFG_CRIT_NRMS && o.criterion != TFG_CRIT_RELRES)
    return fail(TFG_EINVAL, "unknown criterion %d", o.criterion);
  if (o.precision != TFG_FP64 && o.precision != TFG_FP32_IR)
    return fail(TFG_EINVAL, "unknown precision %d", o.precision);
  if (o.deflation < TFG_DEFLATE_AUTO || o.deflation > TFG_DEFLATE_ON)
    return fail(TFG_EINVAL, "unknown deflation mode %d", o.deflation);
  const bool fp32 = o.precision == TFG_FP32_IR;
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->true_relres = -1.0;
  }
  HIP_TRY(hipSetDevice(h->device));
  const size_t N = (size_t)h->nt;
  const double* dval = val;
  const double* db = b;
  if (!device_ptrs) {
    if (!h->val) HIP_TRY(hipMalloc(&h->val, sizeof(double) * (size_t)h->nnz));
    if (!h->b) HIP_TRY(hipMalloc(&h->b, sizeof(double) * N));
    HIP_TRY(hipMemcpyAsync(h->val, val, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->b, b, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->x, x, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
    dval = h->val;
    db = h->b;
  } else {
    HIP_TRY(hipMemcpyAsync(h->x, x, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
  }
  if (fp32) {  // the FP32 form's arrays, once per handle
    const size_t nn = std::max<size_t>(N, 4);
    if (!h->f.val32) HIP_TRY(hipMalloc(&h->f.val32, sizeof(float) * (size_t)std::max(h->nnz, 4)));
    if (!h->f.s64) HIP_TRY(hipMalloc(&h->f.s64, sizeof(double) * nn));
    for (float** v : {&h->f.s32, &h->f.rd32, &h->f.r32, &h->f.p32, &h->f.q32, &h->f.e32})
      if (!*v) HIP_TRY(hipMalloc(v, sizeof(float) * nn));
    if (h->sdg && !h->f.sd32) HIP_TRY(hipMalloc(&h->f.sd32, sizeof(float) * (size_t)h->sd_len));
    if (!h->f.asym) HIP_TRY(hipMalloc(&h->f.asym, sizeof(int) * (size_t)h->nsys));
    HIP_TRY(hipMemsetAsync(h->f.asym, 0, sizeof(int) * (size_t)h->nsys, h->stream));
  }
  { const int rt = batch_tune(h, pr, fp32, o, dval, db, miter, tol, norm); if (rt != TFG_OK) return rt; }
  float ms = 0.f;
  { const int rr = batch_run(h, pr, fp32, o, dval, db, miter, tol, norm, h->cfg[fp32 ? 1 : 0], &ms);
    if (rr != TFG_OK) return rr; }
  const hipStream_t S = h->stream;
  if (!device_ptrs)
    HIP_TRY(hipMemcpyAsync(x, h->x, sizeof(double) * N, hipMemcpyDeviceToHost, S));
  else
    HIP_TRY(hipMemcpyAsync(x, h->x, sizeof(double) * N, hipMemcpyDeviceToDevice, S));
  HIP_TRY(hipStreamSynchronize(S));
  int rate_khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, h->device));
  const std::vector<unsigned long long>& clk = h->clk_host;
  int status = TFG_OK, maxit = 0;
  for (int s = 0; s < h->nsys; ++s) {
    const State& q = h->hst[s];
    niter[s] = q.done ? q.niter : std::max(miter, 0);
    fin_res[s] = q.res;
    maxit = std::max(maxit, niter[s]);
    if (q.status == TFG_EBREAKDOWN) status = TFG_EBREAKDOWN;
    h->done_ms[s] = q.done && rate_khz > 0 ? (double)(clk[1 + s] - clk[0]) / (double)rate_khz : ms;
  }
  if (rep) {
    rep->status = status;
    rep->iterations = maxit;
    rep->solve_ms = ms;
  }
  if (status == TFG_EBREAKDOWN) return fail(TFG_EBREAKDOWN, "Krylov breakdown in a batched system");
  return TFG_OK;
}

int tfg_batch_tuning(tfg_batch* h, int precision, int* chunk, int* diag_classes,
                     int* rows_per_thread, int* probes, double* probe_ms) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (precision != TFG_FP64 && precision != TFG_FP32_IR)
    return fail(TFG_EINVAL, "unknown precision %d", precision);
  const int p = precision == TFG_FP32_IR ? 1 : 0;
  if (chunk) *chunk = !h->tuned[p] ? -1 : (p ? tfg::kBt32Check : h->cfg[p].chunk);
  if (diag_classes) *diag_classes = h->tuned[p] ? h->cfg[p].diag_mask : -1;
  if (rows_per_thread) *rows_per_thread = h->tuned[p] ? h->cfg[p].rows : -1;
  if (probes) *probes = h->probes[p];
  if (probe_ms) *probe_ms = h->probe_ms[p];
  return TFG_OK;
}

int tfg_batch_times(tfg_batch* h, double* ms) {
  if (!h || !ms) return fail(TFG_EINVAL, "null argument");
  for (int s = 0; s < h->nsys; ++s) ms[s] = h->done_ms[s];
  return TFG_OK;
}

int tfg_batch_info(tfg_batch* h, int* nsys, int* n, int* nnz, int* blocks) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (nsys) *nsys = h->nsys;
  if (n) *n = h->nt;
  if (nnz) *nnz = h->nnz;
  if (blocks) *blocks = h->nblk;
  return TFG_OK;
}

}  // extern "C"
