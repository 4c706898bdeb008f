// FP64 kernels of the Jacobi / unpreconditioned CG (and the reference's
// "bicg", which runs the same recurrence) for gfx950.
//
// The reference loop (Sources/Process/Cpu/Native_Mod/Cg.f90:139-224) is
// re-cut into three bandwidth-bound phases per iteration:
//   pupdate : p = z + beta p, z = r * d_inv              (Cg.f90:145,158-171)
//   spmv_pq : q = A p, (p,q) -> alpha = rho / (p,q)       (Cg.f90:176-198)
//   update  : x += alpha p, r -= alpha q, then NRMS sums, x range and the
//             next (z,r) -> res, stop flag, beta          (Cg.f90:204-222 and
//             the rho of the following iteration, Cg.f90:145-156)
// Scalars live in device memory (State).  Every reduction is finished by the
// kernel that consumes it (all of its workgroups sweep the partials in the
// same fixed order), so the host never waits inside the loop and no kernel
// needs atomics or device-scope fences.
//
// Parity: the matrix scaling A*fn of Cg.f90:77-84 is folded into the product
// (each a_ij*fn is rounded exactly as the reference's scaled value), every
// row sum runs in CSR order, and this file is compiled without FP
// contraction, so each SpMV row and each element-wise update is bit-identical
// to the reference's serial arithmetic; only the order of the global dot
// products differs (deterministic trees here).
#include "tfg_spmv.h"
#include <hip/hip_cooperative_groups.h>
#include <algorithm>
#include <type_traits>

namespace tfg {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

// Stopping measure: Normalized_Root_Mean_Square.f90:33-61 (sqrt of the
// diagonal-normalised sum over the x range or over `norm`), or the relative
// residual ||r||_2 / ||b||_2.
__device__ __forceinline__ double measure(const State* st, double nrms, double xmin, double xmax,
                                          double rr) {
  if (st->criterion == TFG_CRIT_RELRES) return sqrt(rr) / st->bnorm;
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(xmax - xmin, kTiny);
  return sqrt(nrms) / range;
}

__device__ __forceinline__ bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ SpmvShape shape_of(const CgArgs& a) {
  SpmvShape m;
  m.n = a.n;
  m.row = a.row;
  m.col = a.col;
  m.blk = a.blk;
  m.nblk = a.nblk;
  m.code = a.code;
  m.tab = a.tab;
  m.ntab = a.ntab;
  m.pat = a.pat;
  m.poff = a.poff;
  m.plen = a.plen;
  m.pslot = a.pslot;
  m.sd = a.sd;
  m.sorder = a.sorder;
  m.nsorder = a.nsorder;
  m.sdg = a.sdg;
  return m;
}

// ---------------------------------------------------------------------------
// Phase kernels
// ---------------------------------------------------------------------------

// sum_i a_ii (partials).                                        Cg.f90:62-73
__global__ __launch_bounds__(kBlock) void k_diag_sum(CgArgs a) {
  __shared__ double s_tmp[4], s_out[1];
  double v[1] = {0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) v[0] = v[0] + a.val[a.row[i]];
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.diag, a.pt.pstride, s_tmp, s_out);
}

// fn = 1/(sum_a/ni) (Cg.f90:75-76), then the initial residual
// r = b*fn - (A*fn) x (Residual_Vector.f90:34-42 on the scaled system), the
// scaled diagonal, and the sums for NRMS(b), NRMS(r), the x range, ||b||,
// ||r|| and rho_1 = (z, r).
template <class E, int PREC>
__global__ __launch_bounds__(kBlock) void k_prologue(CgArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[7 * 4], s_out[7];
  double fn;
  if (a.st->comm) {
    fn = a.st->fn;  // from the global sums (multi-rank solve)
  } else {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.diag, a.pt.pstride, a.pt.g_vec, ops1, s_tmp, s_out);
    const double sum_a = s_out[0] / (double)a.n;
    fn = 1.0 / sum_a;
    if (lead()) a.st->fn = fn;
    __syncthreads();
  }
  double v[7] = {0.0, 0.0, kHuge, -kHuge, 0.0, 0.0, 0.0};
  E::template run<true>(
      sm, shape_of(a), a.val, fn, a.x, s_tmp, s_out, [&](int i) { return a.b[i] * fn; },
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
  publish<7>(v, ops, a.pt.pro, a.pt.pstride, s_tmp, s_out);
}

// One workgroup: the two early exits of Cg.f90:95-123 and rho_1.
__global__ __launch_bounds__(kBlock) void k_start(CgArgs a, int rho_slot) {
  __shared__ double s_tmp[7 * 4], s_out[7];
  const int ops[7] = {OP_SUM, OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM, OP_SUM};
  gather<7>(a.pt.pro, a.pt.pstride, a.pt.g_spmv, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  State* st = a.st;
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
  st->status = TFG_OK;
  st->niter = 0;
  if (bn < st->tol) {                 // Cg.f90:101-104: res stays 0
    st->res = 0.0;
    st->done = 1;
  } else if (res0 < st->tol) {        // Cg.f90:120-123
    st->res = res0;
    st->done = 1;
  } else {
    st->res = res0;
    st->rho[1] = s_out[rho_slot];   // (z, r); (r, r) = (r~0, r) for CGS / BiCGSTAB
    st->done = 0;
  }
}

// Iteration k, first phase.  For k > 1 it finishes the stopping test of
// iteration k-1 (Cg.f90:214-222) and beta = rho_k / rho_{k-1}
// (Cg.f90:165); then p = z + beta p, or p = z when k = 1 (Cg.f90:158-171),
// with z = r * d_inv (Prec_Solve.f90:44-49) recomputed on the fly.
// DX (deferred x, used when the stopping test needs no x range: relres, or
// NRMS with norm): x += alpha_{k-1} p_{k-1} (Cg.f90:206) is done here, on
// the p this kernel reads anyway, instead of in k_update -- the same
// operation on the same operands, one iteration later, 8n bytes fewer per
// iteration.  The last iteration's term is folded by k_fold_x.
// k = 0: the iteration number is State::kcur (graph-replayed chunks).
template <int PREC, bool DX>
__global__ __launch_bounds__(kBlock) void k_pupdate(CgArgs a, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  if (k == 0) k = st->kcur;
  const int stride = gridDim.x * kBlock;
  const int npair = a.n >> 1;
  dbl2* p2 = reinterpret_cast<dbl2*>(a.p);
  dbl2* x2 = reinterpret_cast<dbl2*>(a.x);
  const dbl2* r2 = reinterpret_cast<const dbl2*>(a.r);
  const dbl2* d2 = reinterpret_cast<const dbl2*>(a.ad);
  // this thread's first pair, loaded before the sweep of the partials so
  // that the two memory round trips overlap (small systems are bound by
  // them)
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  const int tp = npair > 0 ? min(t0, npair - 1) : 0;
  dbl2 rv0 = {0.0, 0.0}, dv0 = {1.0, 1.0}, pv0 = {0.0, 0.0}, xv0 = {0.0, 0.0};
  if (npair > 0) {
    rv0 = r2[tp];
    if (PREC == PREC_JACOBI) dv0 = d2[tp];
    pv0 = p2[tp];
    if (DX) xv0 = x2[tp];
  }
  double beta = 0.0;
  if (k > 1) {
    const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
    gather<5>(a.pt.upd, a.pt.pstride, a.pt.g_upd, ops, s_tmp, s_out);
    const double res = measure(st, s_out[0], s_out[1], s_out[2], s_out[4]);
    const bool bad = !isfinite(res);
    if (bad || res < st->tol) {
      if (lead()) {
        st->res = res;
        st->iter = k - 1;
        st->niter = k - 1;
        st->status = bad ? TFG_EBREAKDOWN : TFG_OK;
        st->done = 1;
      }
      return;
    }
    const double rho = s_out[3];
    beta = rho / st->rho[(k - 1) & 1];
    if (lead()) {
      st->rho[k & 1] = rho;
      st->res = res;
      st->iter = k - 1;
      if (DX) st->fold = 0;
    }
  }
  const double alpha = DX && k > 1 ? st->alpha : 0.0;
  auto pair = [&](int t, dbl2 rv, dbl2 dv, dbl2 pv, dbl2 xv) {
    dbl2 z;
    if (PREC == PREC_JACOBI) {
      z.x = rv.x * (1.0 / dv.x);
      z.y = rv.y * (1.0 / dv.y);
    } else {
      z = rv;
    }
    if (k == 1) {
      p2[t] = z;
    } else {
      if (DX) {
        dbl2 xo;
        xo.x = xv.x + alpha * pv.x;
        xo.y = xv.y + alpha * pv.y;
        x2[t] = xo;
      }
      dbl2 o;
      o.x = z.x + beta * pv.x;
      o.y = z.y + beta * pv.y;
      p2[t] = o;
    }
  };
  if (t0 < npair) pair(t0, rv0, dv0, pv0, xv0);
  for (int t = t0 + stride; t < npair; t += stride) {
    const dbl2 rv = r2[t];
    const dbl2 dv = PREC == PREC_JACOBI ? d2[t] : dbl2{1.0, 1.0};
    const dbl2 pv = k == 1 ? dbl2{0.0, 0.0} : p2[t];
    const dbl2 xv = DX && k > 1 ? x2[t] : dbl2{0.0, 0.0};
    pair(t, rv, dv, pv, xv);
  }
  if ((a.n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int i = a.n - 1;
    const double z = PREC == PREC_JACOBI ? a.r[i] * (1.0 / a.ad[i]) : a.r[i];
    if (DX && k > 1) a.x[i] = a.x[i] + alpha * a.p[i];
    a.p[i] = (k == 1) ? z : z + beta * a.p[i];
  }
}

// q = (A*fn) p and the partials of (p, q)                 Cg.f90:176-195
template <class E>
__global__ __launch_bounds__(kBlock) void k_spmv_pq(CgArgs a, int k) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  // graph mode (k = 0): the iteration counter moves on here; no workgroup
  // of this kernel reads it, k_update reads kcur - 1
  if (k == 0 && lead()) a.st->kcur = a.st->kcur + 1;
  const double fn = a.st->fn;
  double v[1] = {0.0};
  E::template run<false>(
      sm, shape_of(a), a.val, fn, a.p, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double qi, double pi, double) {
        a.q[i] = qi;
        v[0] = v[0] + pi * qi;
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

// alpha = rho / (p,q) (Cg.f90:190-198); x += alpha p, r -= alpha q
// (Cg.f90:204-209); partials of the NRMS sum, the x range, the next (z, r)
// and ||r||^2 on the new r and x.
template <int PREC, bool DX>
__global__ __launch_bounds__(kBlock) void k_update(CgArgs a, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  if (k == 0) k = st->kcur - 1;  // the SpMV of this iteration advanced it
  const int stride = gridDim.x * kBlock;
  const int npair = a.n >> 1;
  dbl2* x2 = reinterpret_cast<dbl2*>(a.x);
  dbl2* r2 = reinterpret_cast<dbl2*>(a.r);
  const dbl2* p2 = reinterpret_cast<const dbl2*>(a.p);
  const dbl2* q2 = reinterpret_cast<const dbl2*>(a.q);
  const dbl2* d2 = reinterpret_cast<const dbl2*>(a.ad);
  // this thread's first pair, loaded before the sweep of the (p, q)
  // partials so that the two memory round trips overlap
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  const int tp = npair > 0 ? min(t0, npair - 1) : 0;
  dbl2 rv0 = {0.0, 0.0}, qv0 = {0.0, 0.0}, dv0 = {1.0, 1.0}, xv0 = {0.0, 0.0}, pv0 = {0.0, 0.0};
  if (npair > 0) {
    rv0 = r2[tp];
    qv0 = q2[tp];
    dv0 = d2[tp];
    if (!DX) {
      xv0 = x2[tp];
      pv0 = p2[tp];
    }
  }
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double pq = s_out[0];
  if (pq == 0.0 || !isfinite(pq)) {
    if (lead()) {
      st->status = TFG_EBREAKDOWN;
      st->niter = k;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rho[k & 1] / pq;
  if (DX && lead()) {
    st->alpha = alpha;
    st->fold = 1;
  }
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  if (DX) {
    // r -= alpha q and the sums that need no x (x is folded in k_pupdate),
    // two rows per thread with 16-byte accesses
    auto sums = [&](double rv, double dv) {
      const double z = PREC == PREC_JACOBI ? rv * (1.0 / dv) : rv;
      v[0] = v[0] + (rv * rv) / (dv * dv);
      v[3] = v[3] + z * rv;
      v[4] = v[4] + rv * rv;
    };
    auto pair = [&](int t, dbl2 rv, dbl2 qv, dbl2 dv) {
      dbl2 ro;
      ro.x = rv.x - alpha * qv.x;
      ro.y = rv.y - alpha * qv.y;
      r2[t] = ro;
      sums(ro.x, dv.x);
      sums(ro.y, dv.y);
    };
    if (t0 < npair) pair(t0, rv0, qv0, dv0);
    for (int t = t0 + stride; t < npair; t += stride) pair(t, r2[t], q2[t], d2[t]);
    if ((a.n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      const int i = a.n - 1;
      const double rv = a.r[i] - alpha * a.q[i];
      a.r[i] = rv;
      sums(rv, a.ad[i]);
    }
    const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
    publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
    return;
  }
  auto one = [&](double& xv, double& rv, double pv, double qv, double dv) {
    xv = xv + alpha * pv;
    rv = rv - alpha * qv;
    const double z = PREC == PREC_JACOBI ? rv * (1.0 / dv) : rv;
    v[0] = v[0] + (rv * rv) / (dv * dv);
    v[1] = fmin(v[1], xv);
    v[2] = fmax(v[2], xv);
    v[3] = v[3] + z * rv;
    v[4] = v[4] + rv * rv;
  };
  auto pair = [&](int t, dbl2 xv, dbl2 rv, dbl2 pv, dbl2 qv, dbl2 dv) {
    double x0 = xv.x, x1 = xv.y, r0 = rv.x, r1 = rv.y;
    one(x0, r0, pv.x, qv.x, dv.x);
    one(x1, r1, pv.y, qv.y, dv.y);
    dbl2 xo, ro;
    xo.x = x0;
    xo.y = x1;
    ro.x = r0;
    ro.y = r1;
    x2[t] = xo;
    r2[t] = ro;
  };
  if (t0 < npair) pair(t0, xv0, rv0, pv0, qv0, dv0);
  for (int t = t0 + stride; t < npair; t += stride) pair(t, x2[t], r2[t], p2[t], q2[t], d2[t]);
  if ((a.n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const int i = a.n - 1;
    double xv = a.x[i], rv = a.r[i];
    one(xv, rv, a.p[i], a.q[i], a.ad[i]);
    a.x[i] = xv;
    a.r[i] = rv;
  }
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// One workgroup: stopping test of iteration k (Cg.f90:214-220) so that the
// host-visible state is exact at chunk boundaries and at the end.
// check_omega: BiCGSTAB also stops (breakdown) when omega = 0.
__global__ __launch_bounds__(kBlock) void k_finish(CgArgs a, int k, int check_omega) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  if (k == 0) k = st->kcur - 1;  // the iteration just done
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  gather<5>(a.pt.upd, a.pt.pstride, a.pt.g_upd, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  const double res = measure(st, s_out[0], s_out[1], s_out[2], s_out[4]);
  st->res = res;
  st->iter = k;
  if (!isfinite(res)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
  } else if (res < st->tol) {
    st->niter = k;
    st->done = 1;
  } else if (check_omega && st->omega == 0.0) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k;
    st->done = 1;
  }
}

// The deferred x += alpha_K p_K of the last iteration that updated r.
__global__ __launch_bounds__(kBlock) void k_fold_x(CgArgs a) {
  const State* st = a.st;
  if (!st->fold) return;
  const double alpha = st->alpha;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) a.x[i] = a.x[i] + alpha * a.p[i];
}

// ||b - A x||^2 and ||b||^2 on the unscaled system (FP64), partials.
template <class E>
__global__ __launch_bounds__(kBlock) void k_true_residual(CgArgs a) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[2 * 4], s_out[2];
  double v[2] = {0.0, 0.0};
  E::template run<true>(
      sm, shape_of(a), a.val, 1.0, a.x, s_tmp, s_out, [&](int i) { return a.b[i]; },
      [&](int i, double ri, double, double) {
        const double bi = a.b[i];
        v[0] = v[0] + ri * ri;
        v[1] = v[1] + bi * bi;
      });
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.aux, a.pt.pstride, s_tmp, s_out);
}

// One workgroup: aux0 = ||b - A x|| / ||b||.
__global__ __launch_bounds__(kBlock) void k_true_finish(CgArgs a) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  const int ops[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.aux, a.pt.pstride, a.pt.g_spmv, ops, s_tmp, s_out);
  if (threadIdx.x == 0)
    a.st->aux0 = s_out[1] > 0.0 ? sqrt(s_out[0]) / sqrt(s_out[1]) : sqrt(s_out[0]);
}

// Plain y = A x (RESID: y = b - A x) without scaling or reductions.
template <class E, bool RESID>
__global__ __launch_bounds__(kBlock) void k_spmv_plain(SpmvShape m, const double* __restrict__ val,
                                                       const double* __restrict__ b,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ y) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  E::template run<RESID>(
      sm, m, val, 1.0, x, s_tmp, s_out, [&](int i) { return RESID ? b[i] : 0.0; },
      [&](int i, double yi, double, double) { y[i] = yi; });
}

// q = (A*fn) p and the partials of (p, q) with the symmetric engine (K
// upper diagonals, mode as in SymEngine): the same products as k_spmv_pq.
// Six waves per SIMD leave the pair mode the registers to have all its
// loads in flight at once (at eight it waits for the first ones midway).
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 6)))
void k_spmv_pq_sym(CgArgs a, int k) {
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  if (k == 0 && lead()) a.st->kcur = a.st->kcur + 1;  // as in k_spmv_pq
  const double fn = a.st->fn;
  double v[1] = {0.0};
  SymEngine::run<false, K, MODE>(
      shape_of(a), fn, a.p, [](int) { return 0.0; },
      [&](int i, double qi, double pi, double, bool own) {
        a.q[i] = qi;
        v[0] = v[0] + (own ? pi * qi : 0.0);
      },
      [&](int i, double q0, double q1, double p0, double p1, bool own) {
        dbl2 o;
        o.x = q0;
        o.y = q1;
        *(dbl2*)(a.q + i) = o;
        v[0] = v[0] + (own ? p0 * q0 : 0.0);
        v[0] = v[0] + (own ? p1 * q1 : 0.0);
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

template <int K, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 6)))
void k_spmv_plain_sym(SpmvShape m, const double* __restrict__ x, double* __restrict__ y) {
  SymEngine::run<false, K, MODE>(
      m, 1.0, x, [](int) { return 0.0; },
      [&](int i, double yi, double, double, bool) { y[i] = yi; },
      [&](int i, double y0, double y1, double, double, bool) {
        dbl2 o;
        o.x = y0;
        o.y = y1;
        *(dbl2*)(y + i) = o;
      });
}

// Set-up of the symmetric engine for one val, one thread per row: the
// diagonal and upper entries go to their slots (entries a row lacks stay
// zero from the allocation); every lower entry a(i, j) is compared bit for
// bit with a(j, i), found in row j's pattern (the structure was checked
// symmetric when the handle was made), and any difference sets *flag.
__global__ __launch_bounds__(kBlock) void k_sym_pack(SpmvShape m, const double* __restrict__ val,
                                                     double* __restrict__ sd, int* flag) {
  __shared__ int s_off[kPatCount * kPatMax], s_slot[kPatCount * kPatMax], s_len[kPatCount];
  for (int t = threadIdx.x; t < kPatCount * kPatMax; t += kBlock) {
    s_off[t] = m.poff[t];
    s_slot[t] = m.pslot[t];
  }
  for (int t = threadIdx.x; t < kPatCount; t += kBlock) s_len[t] = m.plen[t];
  __syncthreads();
  const long long S = m.sdg.stride;
  bool bad = false;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m.n; i += stride) {
    const int p = m.pat[i];
    const int L = s_len[p];
    const int rs = m.row[i];
    for (int u = 0; u < L; ++u) {
      const int off = s_off[p * kPatMax + u];
      const double v = val[rs + u];
      if (off >= 0) {
        sd[s_slot[p * kPatMax + u] * S + i] = v;  // diagonal (slot 0) or own upper slot
      } else {
        const int j = i + off, pj = m.pat[j], Lj = s_len[pj];
        int mu = -1;
        for (int t = 0; t < Lj; ++t)
          if (s_off[pj * kPatMax + t] == -off) mu = t;
        if (mu < 0 || __double_as_longlong(val[m.row[j] + mu]) != __double_as_longlong(v)) bad = true;
      }
    }
  }
  if (bad) atomicOr(flag, 1);
}

// Symmetric-engine instantiation for the runtime K: f(integral_constant<K>).
template <typename F>
hipError_t sym_dispatch(int K, F f) {
  switch (K) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Mode of a symmetric product (SymEngine): pairs need 16-byte aligned x, y.
int sym_mode(const SymDiag& g, const void* x, const void* y) {
  if (!g.pair) return 0;
  return (((uintptr_t)x | (uintptr_t)y) % 16) == 0 ? 1 : 2;
}

// ---------------------------------------------------------------------------
// Engine dispatch
// ---------------------------------------------------------------------------
#define TFG_ENGINE_SWITCH(L, CALL) TFG_ENGINE_SWITCH_T(L, double, CALL)

int spmv_blocks_per_cu(int kind) {
  int nb = 0;
  hipError_t e = hipSuccess;
  switch (kind) {
    case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<StreamEngine<kCap, false, double>>, kBlock, 0); break;
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<StreamEngine<kCap2, false, double>>, kBlock, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<WaveEngine<4, double>>, kBlock, 0); break;
    case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<WaveEngine<8, double>>, kBlock, 0); break;
    case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<StreamEngine<kCap, true, double>>, kBlock, 0); break;
    case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<StreamEngine<kCap, true, double, true>>, kBlock, 0); break;
    case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<PatternEngine<kCap, true, double>>, kBlock, 0); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq_sym<3, 1>, kBlock, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmv_pq<WaveEngine<16, double>>, kBlock, 0); break;
  }
  if (e != hipSuccess || nb < 1) return 1;
  return nb;
}

hipError_t launch_diag_sum(const Launch& L, const CgArgs& a) {
  k_diag_sum<<<L.grid_vec, kBlock, 0, L.stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_prologue(const Launch& L, const CgArgs& a, int prec) {
  if (prec == PREC_JACOBI)
    TFG_ENGINE_SWITCH(L, (k_prologue<E, PREC_JACOBI><<<L.grid_spmv, kBlock, 0, L.stream>>>(a)));
  else
    TFG_ENGINE_SWITCH(L, (k_prologue<E, PREC_NONE><<<L.grid_spmv, kBlock, 0, L.stream>>>(a)));
  return hipGetLastError();
}

hipError_t launch_start(const Launch& L, const CgArgs& a, int rho_slot) {
  k_start<<<1, kBlock, 0, L.stream>>>(a, rho_slot);
  return hipGetLastError();
}

hipError_t launch_pupdate(const Launch& L, const CgArgs& a, int prec, int k, bool dx) {
  if (prec == PREC_JACOBI) {
    if (dx) k_pupdate<PREC_JACOBI, true><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
    else k_pupdate<PREC_JACOBI, false><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
  } else {
    if (dx) k_pupdate<PREC_NONE, true><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
    else k_pupdate<PREC_NONE, false><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
  }
  return hipGetLastError();
}

hipError_t launch_fold_x(const Launch& L, const CgArgs& a) {
  k_fold_x<<<L.grid_vec, kBlock, 0, L.stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_spmv_pq(const Launch& L, const CgArgs& a, int k) {
  if (L.sym) {
    const int mode = sym_mode(a.sdg, a.p, a.q);
    return sym_dispatch(a.sdg.k, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if constexpr (K <= kSymPairMax)
        if (mode == 1) {
          k_spmv_pq_sym<K, 1><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, k);
          return;
        }
      if (mode == 2) k_spmv_pq_sym<K, 2><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, k);
      else k_spmv_pq_sym<K, 0><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, k);
    });
  }
  TFG_ENGINE_SWITCH(L, (k_spmv_pq<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, k)));
  return hipGetLastError();
}

hipError_t launch_update(const Launch& L, const CgArgs& a, int prec, int k, bool dx) {
  if (prec == PREC_JACOBI) {
    if (dx) k_update<PREC_JACOBI, true><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
    else k_update<PREC_JACOBI, false><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
  } else {
    if (dx) k_update<PREC_NONE, true><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
    else k_update<PREC_NONE, false><<<L.grid_vec, kBlock, 0, L.stream>>>(a, k);
  }
  return hipGetLastError();
}

hipError_t launch_finish(const Launch& L, const CgArgs& a, int k, int check_omega) {
  k_finish<<<1, kBlock, 0, L.stream>>>(a, k, check_omega);
  return hipGetLastError();
}

hipError_t launch_true_residual(const Launch& L, const CgArgs& a) {
  TFG_ENGINE_SWITCH(L, (k_true_residual<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a)));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_true_finish<<<1, kBlock, 0, L.stream>>>(a);
  return hipGetLastError();
}

// Plain 16-byte streaming copy: a kernel of known traffic (reads n16*16
// bytes, writes n16*16 bytes) used to calibrate the HBM byte counters.
__global__ __launch_bounds__(kBlock) void k_copy16(const dbl2* __restrict__ src,
                                                   dbl2* __restrict__ dst, long long n16) {
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Read-only sweep: every thread sums its 16-byte chunks; one sum per
// workgroup lands in out[blockIdx.x] (so the loads are not dead code).
__global__ __launch_bounds__(kBlock) void k_read16(const dbl2* __restrict__ src,
                                                   double* __restrict__ out, long long n16) {
  __shared__ double s_tmp[4], s_out[1];
  const long long stride = (long long)gridDim.x * kBlock;
  double v[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) {
    const dbl2 c = src[i];
    v[0] = v[0] + (c.x + c.y);
  }
  const int ops[1] = {OP_SUM};
  block_reduce<1>(v, ops, s_tmp, s_out);
  if (threadIdx.x == 0) out[blockIdx.x] = s_out[0];
}

// Write-only fill of 16-byte chunks.
__global__ __launch_bounds__(kBlock) void k_fill16(dbl2* __restrict__ dst, long long n16) {
  const long long stride = (long long)gridDim.x * kBlock;
  const dbl2 z = {1.0, 2.0};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) dst[i] = z;
}

// Zero fill of 16-byte chunks (a kernel, so that no memset node is needed
// where a caller's stream may be captured).
__global__ __launch_bounds__(kBlock) void k_zero16(dbl2* __restrict__ dst, long long n16) {
  const long long stride = (long long)gridDim.x * kBlock;
  const dbl2 z = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) dst[i] = z;
}

hipError_t launch_zero16(hipStream_t stream, void* dst, long long n16) {
  const long long g = (n16 + kBlock - 1) / kBlock;
  const int grid = (int)std::max<long long>(1, std::min<long long>(g, 2048));
  k_zero16<<<grid, kBlock, 0, stream>>>(reinterpret_cast<dbl2*>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_copy16(hipStream_t stream, int grid, const void* src, void* dst, long long n16) {
  k_copy16<<<grid, kBlock, 0, stream>>>(reinterpret_cast<const dbl2*>(src), reinterpret_cast<dbl2*>(dst), n16);
  return hipGetLastError();
}

hipError_t launch_read16(hipStream_t stream, int grid, const void* src, double* out, long long n16) {
  k_read16<<<grid, kBlock, 0, stream>>>(reinterpret_cast<const dbl2*>(src), out, n16);
  return hipGetLastError();
}

hipError_t launch_fill16(hipStream_t stream, int grid, void* dst, long long n16) {
  k_fill16<<<grid, kBlock, 0, stream>>>(reinterpret_cast<dbl2*>(dst), n16);
  return hipGetLastError();
}

// The two ways of separating the phases of an iteration, timed: `reps`
// grid-wide barriers inside one cooperative launch, and `reps` empty
// kernels back to back on one stream (each a kernel boundary).
__global__ __launch_bounds__(kBlock) void k_grid_barriers(int reps, double* sink) {
  cooperative_groups::grid_group g = cooperative_groups::this_grid();
  for (int r = 0; r < reps; ++r) g.sync();
  if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] = (double)reps;
}

__global__ __launch_bounds__(kBlock) void k_boundary(double* sink, int r) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sink[0] = (double)r;
}

hipError_t launch_grid_barriers(hipStream_t stream, int grid, int reps, double* sink) {
  void* args[] = {&reps, &sink};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_grid_barriers), dim3(grid),
                                    dim3(kBlock), args, 0, stream);
}

hipError_t launch_boundary(hipStream_t stream, int grid, double* sink, int r) {
  k_boundary<<<grid, kBlock, 0, stream>>>(sink, r);
  return hipGetLastError();
}

int grid_barrier_fit(int* per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_grid_barriers, kBlock, 0) ==
                 hipSuccess
             ? 1
             : 0;
}

hipError_t launch_spmv_plain(const Launch& L, const SpmvShape& m, const double* val,
                             const double* p, double* q) {
  if (L.sym) {
    const int mode = sym_mode(m.sdg, p, q);
    return sym_dispatch(m.sdg.k, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if constexpr (K <= kSymPairMax)
        if (mode == 1) {
          k_spmv_plain_sym<K, 1><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, p, q);
          return;
        }
      if (mode == 2) k_spmv_plain_sym<K, 2><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, p, q);
      else k_spmv_plain_sym<K, 0><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, p, q);
    });
  }
  TFG_ENGINE_SWITCH(L, (k_spmv_plain<E, false><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, val, nullptr, p, q)));
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_gather(const double* __restrict__ src,
                                                   const int* __restrict__ map,
                                                   double* __restrict__ dst, long long count) {
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < count; k += stride)
    dst[k] = src[map[k]];
}

hipError_t launch_gather(hipStream_t stream, const double* src, const int* map, double* dst,
                         long long count) {
  if (count <= 0) return hipSuccess;
  const long long blocks = (count + kBlock - 1) / kBlock;
  k_gather<<<(int)std::min<long long>(blocks, 4096), kBlock, 0, stream>>>(src, map, dst, count);
  return hipGetLastError();
}

hipError_t launch_sym_pack(hipStream_t stream, int grid, const SpmvShape& m,
                           const double* val, double* sd, int* flag) {
  k_sym_pack<<<grid, kBlock, 0, stream>>>(m, val, sd, flag);
  return hipGetLastError();
}

hipError_t launch_residual_plain(const Launch& L, const SpmvShape& m, const double* val,
                                 const double* b, const double* x, double* r) {
  TFG_ENGINE_SWITCH(L, (k_spmv_plain<E, true><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, val, b, x, r)));
  return hipGetLastError();
}

}  // namespace tfg
