// FP32 Krylov sweeps inside FP64 iterative refinement, for gfx950.
//
// The inner preconditioned CG runs in FP32 (matrix val32 = float(a_ij*fn),
// vectors r32, p32, q32, the correction e32, the product and the Jacobi
// scaling) on the scaled system of Cg.f90:62-84.  Its global dot products are
// accumulated in FP64 and its step lengths alpha, beta are formed in FP64.
// Refinement: whenever the FP32 recursive residual has dropped by the inner
// tolerance since the last refinement, the FP32 correction is folded into
// the FP64 solution (x += e32, k_ir_correct), the true residual r = b - A x
// is recomputed in FP64 (k_prologue), the stopping test is taken on it
// (k_ir_check), it replaces r32 (k_ir_replace) and the inner CG restarts
// from it.
#include "tfg_spmv.h"

namespace tfg {

typedef float flt4 __attribute__((ext_vector_type(4)));

// val32 = float(a_ij * fn) (the product rounded in FP64 first, as the
// FP64 path scales); dinv32 = float(1 / (a_ii * fn)) for Jacobi, else 1.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_convert(CgArgs a, IrArgs f, int nnz) {
  const double fn = a.st->fn;
  float* val32 = const_cast<float*>(f.val32);
  float* dinv32 = const_cast<float*>(f.dinv32);
  const int stride = gridDim.x * kBlock;
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < nnz; j += stride) val32[j] = (float)(a.val[j] * fn);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride)
    dinv32[i] = PREC == PREC_JACOBI ? (float)(1.0 / (a.val[a.row[i]] * fn)) : 1.0f;
}

// One workgroup: the stopping test on the FP64 true residual just computed
// by k_prologue (its seven sums).
__global__ __launch_bounds__(kBlock) void k_ir_check(CgArgs a, int outer) {
  __shared__ double s_tmp[7 * 4], s_out[7];
  const int ops[7] = {OP_SUM, OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM, OP_SUM};
  gather<7>(a.pt.pro, a.pt.pstride, a.pt.g_spmv, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  State* st = a.st;
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(s_out[3] - s_out[2], kTiny);
  if (outer == 0) {
    st->bnorm = sqrt(s_out[4]);
    st->status = TFG_OK;
    st->niter = 0;
    st->done = 0;
    const double bn = st->criterion == TFG_CRIT_RELRES ? (st->bnorm > 0.0 ? 1.0 : 0.0)
                                                       : sqrt(s_out[0]) / range;
    if (bn < st->tol) {  // Cg.f90:101-104
      st->res = 0.0;
      st->done = 1;
      return;
    }
  }
  const double res = st->criterion == TFG_CRIT_RELRES ? sqrt(s_out[5]) / st->bnorm
                                                      : sqrt(s_out[1]) / range;
  st->res = res;
  st->iter = outer;
  if (!isfinite(res)) {
    st->status = TFG_EBREAKDOWN;
    st->done = 1;
  } else if (res < st->tol) {
    st->done = 1;
  }
  st->aux2 = res;
}

// Residual replacement: r32 = float(r) from the FP64 true residual, and the
// partials of (z, r) and ||r32||^2 that the next k_ir_pupdate continues from.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_replace(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  if (a.st->done) return;
  double v[2] = {0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const float r = (float)a.r[i];
    const float z = PREC == PREC_JACOBI ? r * f.dinv32[i] : r;
    f.r32[i] = r;
    v[0] = v[0] + (double)(z * r);
    v[1] = v[1] + (double)r * (double)r;
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// One workgroup: ||r32||^2 after the replacement becomes the reference the
// next refinement is measured against.
__global__ __launch_bounds__(kBlock) void k_ir_replace_finish(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  const int ops[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.upd, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  State* in = f.st_in;
  in->aux1 = s_out[1];  // ||r32||^2 at the refinement
  in->aux0 = 0.0;       // refinement due
  in->done = a.st->done;
  if (!(s_out[1] > 0.0)) in->done = 1;
}

// p32 = z + beta p32 (p32 = z on the first iteration or after a restart),
// beta = rho_k / rho_{k-1} formed in FP64 from the partials of the previous
// update (or replacement).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_pupdate(CgArgs a, IrArgs f, int k, int first) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  State* in = f.st_in;
  if (in->done) return;
  const int ops[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.upd, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  const double rho = s_out[0];
  if (!isfinite(rho) || !isfinite(s_out[1])) {
    if (lead_thread()) {
      in->status = TFG_EBREAKDOWN;
      in->niter = k - 1;
      in->done = 1;
    }
    return;
  }
  const float beta = first ? 0.0f : (float)(rho / in->rho[(k - 1) & 1]);
  if (lead_thread()) in->rho[k & 1] = rho;
  const int stride = gridDim.x * kBlock;
  const int n4 = a.n >> 2;
  flt4* p4 = reinterpret_cast<flt4*>(f.p32);
  const flt4* r4 = reinterpret_cast<const flt4*>(f.r32);
  const flt4* d4 = reinterpret_cast<const flt4*>(f.dinv32);
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n4; t += stride) {
    const flt4 z = PREC == PREC_JACOBI ? r4[t] * d4[t] : r4[t];
    p4[t] = first ? z : z + beta * p4[t];
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < (a.n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
    const float z = PREC == PREC_JACOBI ? f.r32[i] * f.dinv32[i] : f.r32[i];
    f.p32[i] = first ? z : z + beta * f.p32[i];
  }
}

// q32 = A32 p32 (FP32 products and row sums in CSR order) and the FP64
// partials of (p, q).
template <class E>
__global__ __launch_bounds__(kBlock) void k_ir_spmv_pq(CgArgs a, IrArgs f, int k) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  (void)k;
  if (f.st_in->done) return;
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
  double v[1] = {0.0};
  E::template run<false>(
      sm, m, f.val32, 1.0f, f.p32, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, float qi, float pi, float) {
        f.q32[i] = qi;
        v[0] = v[0] + (double)pi * (double)qi;
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

// alpha = rho / (p,q) in FP64; e32 += alpha p32, r32 -= alpha q32; FP64
// partials of the next (z, r) and ||r||^2.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_update(CgArgs a, IrArgs f, int k) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  State* in = f.st_in;
  if (in->done) return;
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double pq = s_out[0];
  if (pq == 0.0 || !isfinite(pq)) {
    if (lead_thread()) {
      in->status = TFG_EBREAKDOWN;
      in->niter = k;
      in->done = 1;
    }
    return;
  }
  const float alpha = (float)(in->rho[k & 1] / pq);
  double v[2] = {0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  const int n4 = a.n >> 2;
  flt4* e4 = reinterpret_cast<flt4*>(f.e32);
  flt4* r4 = reinterpret_cast<flt4*>(f.r32);
  const flt4* p4 = reinterpret_cast<const flt4*>(f.p32);
  const flt4* q4 = reinterpret_cast<const flt4*>(f.q32);
  const flt4* d4 = reinterpret_cast<const flt4*>(f.dinv32);
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n4; t += stride) {
    e4[t] = e4[t] + alpha * p4[t];
    const flt4 r = r4[t] - alpha * q4[t];
    r4[t] = r;
    const flt4 z = PREC == PREC_JACOBI ? r * d4[t] : r;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      v[0] = v[0] + (double)(z[c] * r[c]);
      v[1] = v[1] + (double)r[c] * (double)r[c];
    }
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < (a.n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
    f.e32[i] = f.e32[i] + alpha * f.p32[i];
    const float r = f.r32[i] - alpha * f.q32[i];
    f.r32[i] = r;
    const float z = PREC == PREC_JACOBI ? r * f.dinv32[i] : r;
    v[0] = v[0] + (double)(z * r);
    v[1] = v[1] + (double)r * (double)r;
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// One workgroup at chunk ends: raise the refinement flag once the FP32
// recursive residual has dropped by delta (aux0 = 1).
__global__ __launch_bounds__(kBlock) void k_ir_finish(CgArgs a, IrArgs f, int k, double delta) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  State* in = f.st_in;
  if (in->done) return;
  const int ops[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.upd, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  in->iter = k;
  in->res = sqrt(s_out[1]);
  if (!isfinite(s_out[1])) {
    in->status = TFG_EBREAKDOWN;
    in->niter = k;
    in->done = 1;
  } else if (s_out[1] <= delta * delta * in->aux1) {
    in->aux0 = 1.0;
  }
}

// x += e32 in FP64, e32 = 0 (the FP32 accumulator restarts at every
// refinement, so its rounding stays relative to the increments since then).
__global__ __launch_bounds__(kBlock) void k_ir_correct(CgArgs a, IrArgs f) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    a.x[i] = a.x[i] + (double)f.e32[i];
    f.e32[i] = 0.0f;
  }
}

__global__ __launch_bounds__(kBlock) void k_ir_zero(IrArgs f) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < f.n; i += stride) f.e32[i] = 0.0f;
}

template <class E>
__global__ __launch_bounds__(kBlock) void k_spmv_plain32(SpmvShape m, const float* __restrict__ val,
                                                         const float* __restrict__ x,
                                                         float* __restrict__ y) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  E::template run<false>(sm, m, val, 1.0f, x, s_tmp, s_out, [](int) { return 0.0f; },
                         [&](int i, float yi, float, float) { y[i] = yi; });
}

// ---------------------------------------------------------------------------
hipError_t launch_ir_convert(const Launch& L, const CgArgs& a, const IrArgs& f, int prec,
                             int nnz) {
  if (prec == PREC_JACOBI) k_ir_convert<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, nnz);
  else k_ir_convert<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, nnz);
  return hipGetLastError();
}

hipError_t launch_ir_check(const Launch& L, const CgArgs& a, int outer) {
  k_ir_check<<<1, kBlock, 0, L.stream>>>(a, outer);
  return hipGetLastError();
}

hipError_t launch_ir_replace(const Launch& L, const CgArgs& a, const IrArgs& f, int prec) {
  if (prec == PREC_JACOBI) k_ir_replace<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  else k_ir_replace<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_ir_replace_finish<<<1, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_ir_zero(const Launch& L, const IrArgs& f) {
  k_ir_zero<<<L.grid_vec, kBlock, 0, L.stream>>>(f);
  return hipGetLastError();
}

hipError_t launch_ir_pupdate(const Launch& L, const CgArgs& a, const IrArgs& f, int prec, int k,
                             int first) {
  if (prec == PREC_JACOBI) k_ir_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, k, first);
  else k_ir_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, k, first);
  return hipGetLastError();
}

hipError_t launch_ir_spmv_pq(const Launch& L, const CgArgs& a, const IrArgs& f, int k) {
  TFG_ENGINE_SWITCH_T(L, float, (k_ir_spmv_pq<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, f, k)));
  return hipGetLastError();
}

hipError_t launch_ir_update(const Launch& L, const CgArgs& a, const IrArgs& f, int prec, int k) {
  if (prec == PREC_JACOBI) k_ir_update<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, k);
  else k_ir_update<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f, k);
  return hipGetLastError();
}

hipError_t launch_ir_finish(const Launch& L, const CgArgs& a, const IrArgs& f, int k,
                            double delta) {
  k_ir_finish<<<1, kBlock, 0, L.stream>>>(a, f, k, delta);
  return hipGetLastError();
}

hipError_t launch_ir_correct(const Launch& L, const CgArgs& a, const IrArgs& f) {
  k_ir_correct<<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_spmv_plain32(const Launch& L, const SpmvShape& m, const float* val,
                               const float* p, float* q) {
  TFG_ENGINE_SWITCH_T(L, float, (k_spmv_plain32<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(m, val, p, q)));
  return hipGetLastError();
}

}  // namespace tfg
