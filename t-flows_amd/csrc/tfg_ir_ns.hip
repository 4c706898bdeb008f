// FP32 sweeps of the textbook methods (BiCG, CGS, BiCGSTAB; Barrett et al.,
// "Templates", SIAM 1994, Figs. 2.8-2.10) inside FP64 refinement, for
// gfx950.  No counterpart in the reference (its PETSc path offers the
// methods in FP64 only, Petsc_Interface.cpp:416-436).
//
// The outer FP64 loop (tfg_api.hip, solve_ir_ns) computes the true residual
// r = b - A x and the stopping test in FP64; each inner solve runs one of
// the methods in FP32 on A e = r from e = 0, with the FP32 row-excess
// product of tfg_ir.hip (y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i)), its
// transpose for BiCG (the diagonal slot holding the column sum), FP32
// vectors and FP64 scalars and dot products (each FP32 product exact in
// FP64), until its residual has fallen by the inner tolerance; then
// x += e in FP64.  The kernels follow the FP64 ones in tfg_nonsym.hip step
// for step, with the inner solve's own State (relres criterion, bnorm =
// ||r|| at the restart).
#include "tfg_spmv.h"

namespace tfg {

namespace {

typedef float f32;

__device__ __forceinline__ bool lead2() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ SpmvShape shape2(const CgArgs& a) {
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

__device__ __forceinline__ void stop2(State* st, int k, double res, int status) {
  if (lead2()) {
    st->res = res;
    st->iter = k;
    st->niter = k;
    st->status = status;
    st->done = 1;
  }
}

// Start of inner iteration k: for k > 1 the relres test of iteration k-1 on
// the update partials, then rho_k (as tfg_nonsym.hip's begin_iteration).
__device__ __forceinline__ bool begin32(const CgArgs& a, State* st, int k, bool check_omega,
                                        double* s_tmp, double* s_out, double& rho,
                                        double& rho_old) {
  if (k > 1) {
    const int ops[3] = {OP_SUM, OP_SUM, OP_SUM};
    gather<3>(a.pt.upd, a.pt.pstride, a.pt.g_upd, ops, s_tmp, s_out);
    const double res = sqrt(s_out[2]) / st->bnorm;
    if (!isfinite(res)) { stop2(st, k - 1, res, TFG_EBREAKDOWN); return false; }
    if (res < st->tol) { stop2(st, k - 1, res, TFG_OK); return false; }
    if (check_omega && st->omega == 0.0) { stop2(st, k - 1, res, TFG_EBREAKDOWN); return false; }
    rho = s_out[1];
    rho_old = st->rho[(k - 1) & 1];
    if (lead2()) {
      st->rho[k & 1] = rho;
      st->res = res;
      st->iter = k - 1;
    }
  } else {
    rho = st->rho[1];
    rho_old = 1.0;
  }
  if (rho == 0.0) { stop2(st, k, st->res, TFG_EBREAKDOWN); return false; }
  return true;
}

template <int PREC>
__device__ __forceinline__ f32 mp(f32 v, f32 rd) {
  return PREC == PREC_JACOBI ? v * rd : v;
}

}  // namespace

// Restart of the inner solve from the FP64 true residual a.r: r32 = fl32(r),
// r~ = r32, e = 0, and (r, r), (M^-1 r, r) (BiCG's rho_1) as partials.
__global__ __launch_bounds__(kBlock) void k_ns32_restart(CgArgs a, IrArgs f, NsArgs32 s) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  double v[2] = {0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 r = (f32)a.r[i];
    f.r32[i] = r;
    s.rt[i] = r;
    f.e32[i] = 0.0f;
    v[0] = v[0] + (double)r * (double)r;
    v[1] = v[1] + (double)(f.prec == PREC_JACOBI ? r * f.rd32[i] : r) * (double)r;
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.aux, a.pt.pstride, s_tmp, s_out);
}

// One workgroup: the inner State of the restart (relres against ||r||,
// tol = the inner tolerance, rho_1 of the method).
__global__ __launch_bounds__(kBlock) void k_ns32_begin(CgArgs a, State* in, double tol_in,
                                                       int method) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  const int ops[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.aux, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  in->bnorm = sqrt(s_out[0]);
  in->tol = tol_in;
  in->criterion = TFG_CRIT_RELRES;
  in->res = 1.0;
  in->iter = 0;
  in->niter = 0;
  in->status = TFG_OK;
  in->omega = 1.0;
  in->alpha = 0.0;
  in->rho[1] = method == NS_BICG ? s_out[1] : s_out[0];
  in->done = s_out[0] > 0.0 ? 0 : 1;
}

// ----------------------------------------------------------------- BiCG --
// s.v1 = p~, s.v2 = q~, s.rt = r~
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bicg32_pupdate(CgArgs a, IrArgs f, NsArgs32 s,
                                                           State* in, int k) {
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  double rho, rho_old;
  if (!begin32(a, in, k, false, s_tmp, s_out, rho, rho_old)) return;
  const f32 beta = (f32)(rho / rho_old);
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 d = f.rd32[i];
    const f32 z = mp<PREC>(f.r32[i], d), zt = mp<PREC>(s.rt[i], d);
    if (k == 1) {
      f.p32[i] = z;
      s.v1[i] = zt;
    } else {
      f.p32[i] = z + beta * f.p32[i];
      s.v1[i] = zt + beta * s.v1[i];
    }
  }
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_bicg32_spmv(CgArgs a, IrArgs f, NsArgs32 s,
                                                        State* in) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (in->done) return;
  double v[1] = {0.0};
  E::template run<false, true>(
      sm, shape2(a), f.val32, 1.0f, f.p32, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, f32 qi, f32, f32) {
        f.q32[i] = qi;
        v[0] = v[0] + (double)s.v1[i] * (double)qi;
      });
  E::template run<false, true>(
      sm, shape2(a), s.valt, 1.0f, s.v1, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, f32 qi, f32, f32) { s.v2[i] = qi; });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

// update partials (inner): [0] unused, [1] (M^-1 r, r~) or (r~0, r), [2] (r, r)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bicg32_update(CgArgs a, IrArgs f, NsArgs32 s,
                                                          State* in, int k) {
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop2(in, k, in->res, TFG_EBREAKDOWN); return; }
  const f32 alpha = (f32)(in->rho[k & 1] / den);
  double v[3] = {0.0, 0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 ev = f.e32[i] + alpha * f.p32[i];
    const f32 rv = f.r32[i] - alpha * f.q32[i];
    const f32 tv = s.rt[i] - alpha * s.v2[i];
    f.e32[i] = ev;
    f.r32[i] = rv;
    s.rt[i] = tv;
    v[1] = v[1] + (double)mp<PREC>(rv, f.rd32[i]) * (double)tv;
    v[2] = v[2] + (double)rv * (double)rv;
  }
  const int ops[3] = {OP_SUM, OP_SUM, OP_SUM};
  publish<3>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ------------------------------------------------------------------ CGS --
// s.v1 = u, s.v2 = p^, s.v3 = v^, s.v4 = u^, f.q32 = q, s.rt = r~0
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs32_pupdate(CgArgs a, IrArgs f, NsArgs32 s,
                                                          State* in, int k) {
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  double rho, rho_old;
  if (!begin32(a, in, k, false, s_tmp, s_out, rho, rho_old)) return;
  const f32 beta = (f32)(rho / rho_old);
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    f32 u, p;
    if (k == 1) {
      u = f.r32[i];
      p = u;
    } else {
      const f32 q = f.q32[i];
      u = f.r32[i] + beta * q;
      p = u + beta * (q + beta * f.p32[i]);
    }
    s.v1[i] = u;
    f.p32[i] = p;
    s.v2[i] = mp<PREC>(p, f.rd32[i]);
  }
}

// v = A w and (r~0, v): the first product of CGS and of BiCGSTAB
template <class E>
__global__ __launch_bounds__(kBlock) void k_ns32_spmv_rv(CgArgs a, IrArgs f, State* in,
                                                         const f32* w, f32* out, const f32* rt) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (in->done) return;
  double v[1] = {0.0};
  E::template run<false, true>(
      sm, shape2(a), f.val32, 1.0f, w, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, f32 qi, f32, f32) {
        out[i] = qi;
        v[0] = v[0] + (double)rt[i] * (double)qi;
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cgs32_mid(CgArgs a, IrArgs f, NsArgs32 s, State* in,
                                                      int k) {
  __shared__ double s_tmp[4], s_out[1];
  if (in->done) return;
  const int ops1[1] = {OP_SUM};
  gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop2(in, k, in->res, TFG_EBREAKDOWN); return; }
  const double alpha = in->rho[k & 1] / den;
  if (lead2()) in->alpha = alpha;
  const f32 a32 = (f32)alpha;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 u = s.v1[i];
    const f32 q = u - a32 * s.v3[i];
    const f32 uh = mp<PREC>(u + q, f.rd32[i]);
    f.q32[i] = q;
    s.v4[i] = uh;
    f.e32[i] = f.e32[i] + a32 * uh;
  }
}

// q^ = A u^ and r -= alpha q^, with the stopping and rho sums (on the SpMV
// grid: the host sets g_upd)
template <class E>
__global__ __launch_bounds__(kBlock) void k_cgs32_spmv_update(CgArgs a, IrArgs f, NsArgs32 s,
                                                              State* in) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  const f32 alpha = (f32)in->alpha;
  double v[3] = {0.0, 0.0, 0.0};
  E::template run<false, true>(
      sm, shape2(a), f.val32, 1.0f, s.v4, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, f32 qh, f32, f32) {
        const f32 rv = f.r32[i] - alpha * qh;
        f.r32[i] = rv;
        v[1] = v[1] + (double)s.rt[i] * (double)rv;
        v[2] = v[2] + (double)rv * (double)rv;
      });
  const int ops[3] = {OP_SUM, OP_SUM, OP_SUM};
  publish<3>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ------------------------------------------------------------- BiCGSTAB --
// f.q32 = v, s.v1 = s, s.v2 = p^, s.v3 = s^, s.v4 = t, s.rt = r~0
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs32_pupdate(CgArgs a, IrArgs f, NsArgs32 s,
                                                           State* in, int k) {
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  double rho, rho_old;
  if (!begin32(a, in, k, true, s_tmp, s_out, rho, rho_old)) return;
  const double omega = in->omega;
  const f32 beta = (f32)((rho / rho_old) * (in->alpha / omega));
  const f32 om = (f32)omega;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 p = k == 1 ? f.r32[i] : f.r32[i] + beta * (f.p32[i] - om * f.q32[i]);
    f.p32[i] = p;
    s.v2[i] = mp<PREC>(p, f.rd32[i]);
  }
}

// alpha; s = r - alpha v, s^ = M^-1 s, and (s, s) for the early exit (ext row 3)
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs32_mid(CgArgs a, IrArgs f, NsArgs32 s, State* in,
                                                       int k) {
  __shared__ double s_tmp[4], s_out[1];
  if (in->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double den = s_out[0];
  if (den == 0.0 || !isfinite(den)) { stop2(in, k, in->res, TFG_EBREAKDOWN); return; }
  const double alpha = in->rho[k & 1] / den;
  if (lead2()) in->alpha = alpha;
  const f32 a32 = (f32)alpha;
  double v[1] = {0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 sv = f.r32[i] - a32 * f.q32[i];
    s.v1[i] = sv;
    s.v3[i] = mp<PREC>(sv, f.rd32[i]);
    v[0] = v[0] + (double)sv * (double)sv;
  }
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.ext + 3 * (size_t)a.pt.pstride, a.pt.pstride, s_tmp, s_out);
}

// the early-exit test on s; otherwise t = A s^ with (t, s) and (t, t)
template <class E>
__global__ __launch_bounds__(kBlock) void k_bcgs32_spmv_t(CgArgs a, IrArgs f, NsArgs32 s,
                                                          State* in) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[2 * 4], s_out[2];
  if (in->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.ext + 3 * (size_t)a.pt.pstride, a.pt.pstride, a.pt.g_vec, ops1, s_tmp, s_out);
  }
  const double res = sqrt(s_out[0]) / in->bnorm;
  if (!isfinite(res) || res < in->tol) return;  // decided by k_bcgs32_update
  __syncthreads();
  double v[2] = {0.0, 0.0};
  E::template run<false, true>(
      sm, shape2(a), f.val32, 1.0f, s.v3, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, f32 ti, f32, f32) {
        s.v4[i] = ti;
        v[0] = v[0] + (double)ti * (double)s.v1[i];
        v[1] = v[1] + (double)ti * (double)ti;
      });
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.ext + 4 * (size_t)a.pt.pstride, a.pt.pstride, s_tmp, s_out);
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_bcgs32_update(CgArgs a, IrArgs f, NsArgs32 s,
                                                          State* in, int k) {
  __shared__ double s_tmp[3 * 4], s_out[3];
  if (in->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.ext + 3 * (size_t)a.pt.pstride, a.pt.pstride, a.pt.g_vec, ops1, s_tmp, s_out);
  }
  const double res_s = sqrt(s_out[0]) / in->bnorm;
  const f32 alpha = (f32)in->alpha;
  const int stride = gridDim.x * kBlock;
  if (!isfinite(res_s)) { stop2(in, k, res_s, TFG_EBREAKDOWN); return; }
  if (res_s < in->tol) {
    // early exit on s: e += alpha p^, r = s
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
      f.e32[i] = f.e32[i] + alpha * s.v2[i];
      f.r32[i] = s.v1[i];
    }
    stop2(in, k, res_s, TFG_OK);
    return;
  }
  __syncthreads();
  {
    const int ops2[2] = {OP_SUM, OP_SUM};
    gather<2>(a.pt.ext + 4 * (size_t)a.pt.pstride, a.pt.pstride, a.pt.g_spmv, ops2, s_tmp, s_out);
  }
  const double ts = s_out[0], tt = s_out[1];
  if (tt == 0.0 || !isfinite(tt)) { stop2(in, k, in->res, TFG_EBREAKDOWN); return; }
  const double omega = ts / tt;
  if (lead2()) in->omega = omega;
  const f32 om = (f32)omega;
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const f32 eh = f.e32[i] + alpha * s.v2[i];
    const f32 ev = eh + om * s.v3[i];
    const f32 rv = s.v1[i] - om * s.v4[i];
    f.e32[i] = ev;
    f.r32[i] = rv;
    v[1] = v[1] + (double)s.rt[i] * (double)rv;
    v[2] = v[2] + (double)rv * (double)rv;
  }
  const int ops[3] = {OP_SUM, OP_SUM, OP_SUM};
  publish<3>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// --------------------------------------------------------- refinement ----
// x += e (FP64), then the FP64 true residual and its stopping sums come from
// k_ir_resid (State::refine is set for the whole solve) and the test from
// k_ns32_test.
__global__ __launch_bounds__(kBlock) void k_ns32_fold(CgArgs a, IrArgs f) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride)
    a.x[i] = a.x[i] + (double)f.e32[i];
}

// One workgroup: the stopping test on the FP64 true residual (k_ir_resid's
// partials); niter = the FP32 iterations so far.
__global__ __launch_bounds__(kBlock) void k_ns32_test(CgArgs a, int iters, int miter) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  gather<5>(a.pt.ext, a.pt.pstride, a.pt.g_spmv, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  State* st = a.st;
  double res;
  if (st->criterion == TFG_CRIT_RELRES) {
    res = sqrt(s_out[3]) / st->bnorm;
  } else {
    const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(s_out[2] - s_out[1], kTiny);
    res = sqrt(s_out[0]) / range;
  }
  st->res = res;
  st->iter = iters;
  st->nrefine = st->nrefine + 1;
  if (!isfinite(res)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = iters;
    st->done = 1;
  } else if (res < st->tol || iters >= miter) {
    st->niter = iters;
    st->done = 1;
  }
}

// FP32 transpose of the row-excess matrix: off the diagonal
// fl32(a_ji fn) (the entry at tpos), in the diagonal slot the compensated
// column sum s^T_i = sum_j a_ji fn.
__global__ __launch_bounds__(kBlock) void k_ns32_valt(CgArgs a, const int* __restrict__ tpos,
                                                      f32* __restrict__ valt) {
  const double fn = a.st->fn;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const int rs = a.row[i], re = a.row[i + 1];
    double s = 0.0, c = 0.0;
    for (int j = rs; j < re; ++j) {
      const double v = a.val[tpos[j]] * fn;
      const double t = s + v;
      if (fabs(s) >= fabs(v)) c = c + ((s - t) + v);
      else c = c + ((v - t) + s);
      s = t;
      if (j != rs) valt[j] = (f32)v;
    }
    valt[rs] = (f32)(s + c);
  }
}

// ------------------------------------------------------------- launchers --
#define TFG_NS32_SWITCH(L, CALL)                                                 \
  do {                                                                           \
    switch ((L).kind) {                                                          \
      case K_STAGE2048:                                                          \
        { using E = StreamEngine<kCap, false, float>; CALL; } break;             \
      case K_PIPE2048_CODED:                                                     \
        { using E = StreamEngine<kCap, true, float, true>; CALL; } break;        \
      case K_PIPE_PATTERN:                                                       \
        { using E = PatternEngine<kCap, true, float>; CALL; } break;             \
      default:                                                                   \
        if ((L).vec_val) { using E = StreamEngine<kCap, true, float>; CALL; }    \
        else { using E = ScalarStage<kCap, float>; CALL; }                       \
        break;                                                                   \
    }                                                                            \
  } while (0)

hipError_t launch_ns32_valt(const Launch& L, const CgArgs& a, const int* tpos, float* valt) {
  k_ns32_valt<<<L.grid_vec, kBlock, 0, L.stream>>>(a, tpos, valt);
  return hipGetLastError();
}

hipError_t launch_ns32_restart(const Launch& L, const CgArgs& a, const IrArgs& f,
                               const NsArgs32& s, State* in, double tol_in, int method) {
  k_ns32_restart<<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ns32_begin<<<1, kBlock, 0, L.stream>>>(a, in, tol_in, method);
  return hipGetLastError();
}

hipError_t launch_ns32_iteration(const Launch& L, const Launch& L32, const CgArgs& a,
                                 const IrArgs& f, const NsArgs32& s, State* in, int method,
                                 int k, hipEvent_t ev0, hipEvent_t ev1) {
  const bool jac = f.prec == PREC_JACOBI;
  hipError_t e;
#define TFG_TRY_LAUNCH(stmt) do { stmt; e = hipGetLastError(); if (e != hipSuccess) return e; } while (0)
#define TFG_EV(ev) do { if (ev) { e = hipEventRecord(ev, L.stream); if (e != hipSuccess) return e; } } while (0)
  const int gv = L.grid_vec, gs = L32.grid_spmv;
  switch (method) {
    case NS_BICG:
      if (jac) TFG_TRY_LAUNCH((k_bicg32_pupdate<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_bicg32_pupdate<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_NS32_SWITCH(L32, (k_bicg32_spmv<E><<<gs, kBlock, 0, L.stream>>>(a, f, s, in))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_bicg32_update<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_bicg32_update<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      break;
    case NS_CGS:
      if (jac) TFG_TRY_LAUNCH((k_cgs32_pupdate<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_cgs32_pupdate<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_NS32_SWITCH(L32, (k_ns32_spmv_rv<E><<<gs, kBlock, 0, L.stream>>>(a, f, in, s.v2, s.v3, s.rt))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_cgs32_mid<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_cgs32_mid<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      TFG_TRY_LAUNCH(TFG_NS32_SWITCH(L32, (k_cgs32_spmv_update<E><<<gs, kBlock, 0, L.stream>>>(a, f, s, in))));
      break;
    default:  // NS_BCGS
      if (jac) TFG_TRY_LAUNCH((k_bcgs32_pupdate<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_bcgs32_pupdate<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      TFG_EV(ev0);
      TFG_TRY_LAUNCH(TFG_NS32_SWITCH(L32, (k_ns32_spmv_rv<E><<<gs, kBlock, 0, L.stream>>>(a, f, in, s.v2, f.q32, s.rt))));
      TFG_EV(ev1);
      if (jac) TFG_TRY_LAUNCH((k_bcgs32_mid<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_bcgs32_mid<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      TFG_TRY_LAUNCH(TFG_NS32_SWITCH(L32, (k_bcgs32_spmv_t<E><<<gs, kBlock, 0, L.stream>>>(a, f, s, in))));
      if (jac) TFG_TRY_LAUNCH((k_bcgs32_update<PREC_JACOBI><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      else TFG_TRY_LAUNCH((k_bcgs32_update<PREC_NONE><<<gv, kBlock, 0, L.stream>>>(a, f, s, in, k)));
      break;
  }
#undef TFG_TRY_LAUNCH
#undef TFG_EV
  return hipSuccess;
}

hipError_t launch_ns32_fold_test(const Launch& L, const CgArgs& a, const IrArgs& f, int iters,
                                 int miter) {
  k_ns32_fold<<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_ir_resid(L, a, f);
  if (e != hipSuccess) return e;
  k_ns32_test<<<1, kBlock, 0, L.stream>>>(a, iters, miter);
  return hipGetLastError();
}

}  // namespace tfg
