// FP64 kernels of the textbook methods for non-symmetric systems: BiCG,
// CGS and BiCGSTAB (Barrett et al., "Templates for the Solution of Linear
// Systems", SIAM 1994, Figs. 2.8-2.10), the methods T-Flows' PETSc path
// offers by name (KSPBICG, KSPCGS, KSPBCGS; Sources/Process/Cpu/
// Petsc_Interface.cpp:416-436).  They share the CG path's prologue
// (normalisation folded into the products, r0, the early exits), stopping
// measure and conventions (Cg.f90:62-123, 214-228), and its reduction
// scheme: every phase publishes per-workgroup partials and the next phase
// sweeps them in a fixed order.  The shadow residual is r~0 = r0.
//
// Per iteration (Jacobi M applied as r * (1/a_ii), exactly as Prec_Solve):
//   BiCG     pupdate (p, p~)   | spmv q = A p, q~ = A^T p~, (p~,q) | update
//   CGS      pupdate (u, p, p^)| spmv v^ = A p^, (r~0,v^) | mid (q, u^, x)
//            | spmv q^ = A u^ fused with r -= alpha q^ and the sums
//   BiCGSTAB pupdate (p, p^)   | spmv v = A p^, (r~0,v) | mid (s, s^, the
//            stopping test on s) | spmv t = A s^, (t,s), (t,t) | update
// A^T p~ runs row by row over the same (symmetric) pattern with the
// transposed values val(tpos(j)), so every product row is a CSR-order sum
// like the reference's and bit-identical to the oracle's.
#include "tfg_spmv.h"

namespace tfg {

namespace {

__device__ __forceinline__ bool lead1() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ double measure1(const State* st, double nrms, double xmin,
                                           double xmax, double rr) {
  if (st->criterion == TFG_CRIT_RELRES) return sqrt(rr) / st->bnorm;
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(xmax - xmin, kTiny);
  return sqrt(nrms) / range;
}

__device__ __forceinline__ SpmvShape shape1(const CgArgs& a) {
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
  return m;
}

__device__ __forceinline__ void stop_at(State* st, int k, double res, int status) {
  if (lead1()) {
    st->res = res;
    st->iter = k;
    st->niter = k;
    st->status = status;
    st->done = 1;
  }
}

// Start of iteration k: for k > 1 the stopping test of iteration k-1 on the
// update partials (Cg.f90:214-220), then rho_k.  Returns false when the solve
// stops here.  Every workgroup evaluates the same partials, so all agree.
__device__ __forceinline__ bool begin_iteration(const CgArgs& a, int k, bool check_omega,
                                                double* s_tmp, double* s_out, double& rho,
                                                double& rho_old) {
  State* st = a.st;
  if (k > 1) {
    const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
    gather<5>(a.pt.upd, a.pt.pstride, a.pt.g_upd, ops, s_tmp, s_out);
    const double res = measure1(st, s_out[0], s_out[1], s_out[2], s_out[4]);
    if (!isfinite(res)) { stop_at(st, k - 1, res, TFG_EBREAKDOWN); return false; }
    if (res < st->tol) { stop_at(st, k - 1, res, TFG_OK); return false; }
    if (check_omega && st->omega == 0.0) { stop_at(st, k - 1, res, TFG_EBREAKDOWN); return false; }
    rho = s_out[3];
    rho_old = st->rho[(k - 1) & 1];
    if (lead1()) {
      st->rho[k & 1] = rho;
      st->res = res;
      st->iter = k - 1;
    }
  } else {
    rho = st->rho[1];
    rho_old = 1.0;
  }
  if (rho == 0.0) { stop_at(st, k, st->res, TFG_EBREAKDOWN); return false; }
  return true;
}

template <int PREC>
__device__ __forceinline__ double mprec(double v, double d) {
  return PREC == PREC_JACOBI ? v * (1.0 / d) : v;
}

}  // namespace

// r~ = r0 (the shadow residual of all three methods).
__global__ __launch_bounds__(kBlock) void k_ns_init(CgArgs a, NsArgs s) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) s.rt[i] = a.r[i];
}

// tpos[j] = position of the transposed entry of entry j; *bad = 1 when the
// pattern is not symmetric.  One thread per row; rows are short.
__global__ __launch_bounds__(kBlock) void k_transpose_pos(int n, const int* __restrict__ row,
                                                          const int* __restrict__ col,
                                                          int* __restrict__ tpos, int* bad) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  for (int j = row[i]; j < row[i + 1]; ++j) {
    const int c = col[j];
    int f = -1;
    for (int m = row[c]; m < row[c + 1]; ++m)
      if (col[m] == i) { f = m; break; }
    tpos[j] = f < 0 ? j : f;
    if (f < 0) bad[0] = 1;
  }
}

// valt[j] = val[tpos[j]] (unscaled; the product scales like A's).
__global__ __launch_bounds__(kBlock) void k_gather_valt(long long nnz, const int* __restrict__ tpos,
                                                        const double* __restrict__ val,
                                                        double* __restrict__ valt) {
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < nnz; j += stride)
    valt[j] = val[tpos[j]];
}

// ----------------------------------------------------------------- BiCG --
// s.v1 = p~, s.v2 = q~, s.rt = r~
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bicg_pupdate(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  if (a.st->done) return;
  double rho, rho_old;
  if (!begin_iteration(a, k, false, s_tmp, s_out, rho, rho_old)) return;
  const double beta = rho / rho_old;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double d = a.ad[i];
    const double z = mprec<PREC>(a.r[i], d), zt = mprec<PREC>(s.rt[i], d);
    if (k == 1) {
      a.p[i] = z;
      s.v1[i] = zt;
    } else {
      a.p[i] = z + beta * a.p[i];
      s.v1[i] = zt + beta * s.v1[i];
    }
  }
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_bicg_spmv(CgArgs a, NsArgs s) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  const double fn = a.st->fn;
  double v[1] = {0.0};
  E::template run<false>(
      sm, shape1(a), a.val, fn, a.p, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double qi, double, double) {
        a.q[i] = qi;
        v[0] = v[0] + s.v1[i] * qi;
      });
  E::template run<false>(
      sm, shape1(a), s.valt, fn, s.v1, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double qi, double, double) { s.v2[i] = qi; });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bicg_update(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop_at(st, k, st->res, TFG_EBREAKDOWN); return; }
  const double alpha = st->rho[k & 1] / den;
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double d = a.ad[i];
    const double xv = a.x[i] + alpha * a.p[i];
    const double rv = a.r[i] - alpha * a.q[i];
    const double tv = s.rt[i] - alpha * s.v2[i];
    a.x[i] = xv;
    a.r[i] = rv;
    s.rt[i] = tv;
    v[0] = v[0] + (rv * rv) / (d * d);
    v[1] = fmin(v[1], xv);
    v[2] = fmax(v[2], xv);
    v[3] = v[3] + mprec<PREC>(rv, d) * tv;
    v[4] = v[4] + rv * rv;
  }
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ------------------------------------------------------------------ CGS --
// s.v1 = u, s.v2 = p^, s.v3 = v^, s.v4 = u^, a.q = q, s.rt = r~0
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs_pupdate(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  if (a.st->done) return;
  double rho, rho_old;
  if (!begin_iteration(a, k, false, s_tmp, s_out, rho, rho_old)) return;
  const double beta = rho / rho_old;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    double u, p;
    if (k == 1) {
      u = a.r[i];
      p = u;
    } else {
      const double q = a.q[i];
      u = a.r[i] + beta * q;
      p = u + beta * (q + beta * a.p[i]);
    }
    s.v1[i] = u;
    a.p[i] = p;
    s.v2[i] = mprec<PREC>(p, a.ad[i]);
  }
}

// v = A w and (r~0, v): the first product of CGS and of BiCGSTAB
template <class E>
__global__ __launch_bounds__(kBlock) void k_ns_spmv_rv(CgArgs a, const double* w, double* out,
                                                       const double* rt) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  const double fn = a.st->fn;
  double v[1] = {0.0};
  E::template run<false>(
      sm, shape1(a), a.val, fn, w, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double qi, double, double) {
        out[i] = qi;
        v[0] = v[0] + rt[i] * qi;
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs_mid(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[4], s_out[1];
  State* st = a.st;
  if (st->done) return;
  const int ops1[1] = {OP_SUM};
  gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop_at(st, k, st->res, TFG_EBREAKDOWN); return; }
  const double alpha = st->rho[k & 1] / den;
  if (lead1()) st->alpha = alpha;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double u = s.v1[i];
    const double q = u - alpha * s.v3[i];
    const double uh = mprec<PREC>(u + q, a.ad[i]);
    a.q[i] = q;
    s.v4[i] = uh;
    a.x[i] = a.x[i] + alpha * uh;
  }
}

// q^ = A u^ and r -= alpha q^, with the stopping and rho sums (this kernel
// produces the update partials on the SpMV grid: the host sets g_upd)
template <class E>
__global__ __launch_bounds__(kBlock) void k_cgs_spmv_update(CgArgs a, NsArgs s) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  const double fn = st->fn, alpha = st->alpha;
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  E::template run<false>(
      sm, shape1(a), a.val, fn, s.v4, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double qh, double, double d) {
        const double rv = a.r[i] - alpha * qh;
        const double xv = a.x[i];
        a.r[i] = rv;
        v[0] = v[0] + (rv * rv) / (d * d);
        v[1] = fmin(v[1], xv);
        v[2] = fmax(v[2], xv);
        v[3] = v[3] + s.rt[i] * rv;
        v[4] = v[4] + rv * rv;
      });
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ------------------------------------------------------------- BiCGSTAB --
// a.q = v, s.v1 = s, s.v2 = p^, s.v3 = s^, s.v4 = t, s.rt = r~0
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs_pupdate(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  if (a.st->done) return;
  double rho, rho_old;
  if (!begin_iteration(a, k, true, s_tmp, s_out, rho, rho_old)) return;
  const double omega = a.st->omega;
  const double beta = (rho / rho_old) * (a.st->alpha / omega);
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double p = k == 1 ? a.r[i] : a.r[i] + beta * (a.p[i] - omega * a.q[i]);
    a.p[i] = p;
    s.v2[i] = mprec<PREC>(p, a.ad[i]);
  }
}

// alpha; s = r - alpha v, s^ = M^-1 s, and the stopping sums of s with the
// x + alpha p^ it would give (ext rows 0..3)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs_mid(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[4 * 4], s_out[4];
  State* st = a.st;
  if (st->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop_at(st, k, st->res, TFG_EBREAKDOWN); return; }
  const double alpha = st->rho[k & 1] / den;
  if (lead1()) st->alpha = alpha;
  double v[4] = {0.0, kHuge, -kHuge, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double d = a.ad[i];
    const double sv = a.r[i] - alpha * a.q[i];
    const double xh = a.x[i] + alpha * s.v2[i];
    s.v1[i] = sv;
    s.v3[i] = mprec<PREC>(sv, d);
    v[0] = v[0] + (sv * sv) / (d * d);
    v[1] = fmin(v[1], xh);
    v[2] = fmax(v[2], xh);
    v[3] = v[3] + sv * sv;
  }
  const int ops[4] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM};
  publish<4>(v, ops, a.pt.ext, a.pt.pstride, s_tmp, s_out);
}

// the early-exit test on s; otherwise t = A s^ with (t, s) and (t, t)
template <class E>
__global__ __launch_bounds__(kBlock) void k_bcgs_spmv_t(CgArgs a, NsArgs s) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4 * 4], s_out[4];
  State* st = a.st;
  if (st->done) return;
  {
    const int ops4[4] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM};
    gather<4>(a.pt.ext, a.pt.pstride, a.pt.g_vec, ops4, s_tmp, s_out);
  }
  const double res = measure1(st, s_out[0], s_out[1], s_out[2], s_out[3]);
  if (!isfinite(res) || res < st->tol) return;  // decided by k_bcgs_update
  __syncthreads();
  const double fn = st->fn;
  double v[2] = {0.0, 0.0};
  E::template run<false>(
      sm, shape1(a), a.val, fn, s.v3, s_tmp, s_out, [](int) { return 0.0; },
      [&](int i, double ti, double, double) {
        s.v4[i] = ti;
        v[0] = v[0] + ti * s.v1[i];
        v[1] = v[1] + ti * ti;
      });
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.ext + 4 * (size_t)a.pt.pstride, a.pt.pstride, s_tmp, s_out);
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs_update(CgArgs a, NsArgs s, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  const int ops4[4] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM};
  gather<4>(a.pt.ext, a.pt.pstride, a.pt.g_vec, ops4, s_tmp, s_out);
  const double res_s = measure1(st, s_out[0], s_out[1], s_out[2], s_out[3]);
  const double alpha = st->alpha;
  const int stride = gridDim.x * kBlock;
  if (!isfinite(res_s)) { stop_at(st, k, res_s, TFG_EBREAKDOWN); return; }
  if (res_s < st->tol) {
    // early exit on s: x += alpha p^, r = s
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
      a.x[i] = a.x[i] + alpha * s.v2[i];
      a.r[i] = s.v1[i];
    }
    stop_at(st, k, res_s, TFG_OK);
    return;
  }
  __syncthreads();
  {
    const int ops2[2] = {OP_SUM, OP_SUM};
    gather<2>(a.pt.ext + 4 * (size_t)a.pt.pstride, a.pt.pstride, a.pt.g_spmv, ops2, s_tmp, s_out);
  }
  const double ts = s_out[0], tt = s_out[1];
  if (tt == 0.0 || !isfinite(tt)) { stop_at(st, k, st->res, TFG_EBREAKDOWN); return; }
  const double omega = ts / tt;
  if (lead1()) st->omega = omega;
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double d = a.ad[i];
    const double xh = a.x[i] + alpha * s.v2[i];
    const double xv = xh + omega * s.v3[i];
    const double rv = s.v1[i] - omega * s.v4[i];
    a.x[i] = xv;
    a.r[i] = rv;
    v[0] = v[0] + (rv * rv) / (d * d);
    v[1] = fmin(v[1], xv);
    v[2] = fmax(v[2], xv);
    v[3] = v[3] + s.rt[i] * rv;
    v[4] = v[4] + rv * rv;
  }
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ------------------------------------------------------------- launchers --
#define TFG_ENGINE_SWITCH_NS(L, CALL) TFG_ENGINE_SWITCH_T(L, double, CALL)

hipError_t launch_ns_init(const Launch& L, const CgArgs& a, const NsArgs& s) {
  k_ns_init<<<L.grid_vec, kBlock, 0, L.stream>>>(a, s);
  return hipGetLastError();
}

hipError_t launch_transpose_pos(hipStream_t stream, int n, const int* row, const int* col,
                                int* tpos, int* bad) {
  k_transpose_pos<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(n, row, col, tpos, bad);
  return hipGetLastError();
}

hipError_t launch_gather_valt(const Launch& L, long long nnz, const int* tpos, const double* val,
                              double* valt) {
  k_gather_valt<<<L.grid_vec, kBlock, 0, L.stream>>>(nnz, tpos, val, valt);
  return hipGetLastError();
}

hipError_t launch_ns_iteration(const Launch& L, const CgArgs& a, const NsArgs& s, int method,
                               int prec, int k, hipEvent_t ev0, hipEvent_t ev1) {
  const bool jac = prec == PREC_JACOBI;
  hipError_t e;
#define TFG_TRY_LAUNCH(stmt) do { stmt; e = hipGetLastError(); if (e != hipSuccess) return e; } while (0)
#define TFG_EV(ev) do { if (ev) { e = hipEventRecord(ev, L.stream); if (e != hipSuccess) return e; } } while (0)
  switch (method) {
    case NS_BICG:
      if (jac) TFG_TRY_LAUNCH((k_bicg_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_bicg_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_ENGINE_SWITCH_NS(L, (k_bicg_spmv<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, s))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_bicg_update<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_bicg_update<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      break;
    case NS_CGS:
      if (jac) TFG_TRY_LAUNCH((k_cgs_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_cgs_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_ENGINE_SWITCH_NS(L, (k_ns_spmv_rv<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, s.v2, s.v3, s.rt))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_cgs_mid<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_cgs_mid<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      TFG_TRY_LAUNCH(TFG_ENGINE_SWITCH_NS(L, (k_cgs_spmv_update<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, s))));
      break;
    default:  // NS_BCGS
      if (jac) TFG_TRY_LAUNCH((k_bcgs_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_bcgs_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_ENGINE_SWITCH_NS(L, (k_ns_spmv_rv<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, s.v2, a.q, s.rt))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_bcgs_mid<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_bcgs_mid<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      TFG_TRY_LAUNCH(TFG_ENGINE_SWITCH_NS(L, (k_bcgs_spmv_t<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, s))));
      if (jac) TFG_TRY_LAUNCH((k_bcgs_update<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      else TFG_TRY_LAUNCH((k_bcgs_update<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, s, k)));
      break;
  }
#undef TFG_TRY_LAUNCH
#undef TFG_EV
  return hipSuccess;
}

}  // namespace tfg
