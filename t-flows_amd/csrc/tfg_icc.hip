// Incomplete Cholesky (DIC) preconditioner of Native_Mod on the GPU:
// Prec_Form.f90:68-78 and Prec_Solve.f90:64-92 (the reference's default
// "icc"), exact to the bit.
//
// Both recurrences run over rows in increasing order (form, forward
// substitution) or decreasing order (backward substitution), and row i needs
// the finished values of its lower (upper) neighbours only.  Rows are
// therefore grouped in dependency levels computed once per structure
// (level(i) = 1 + the largest level among the neighbours it reads), and one
// launch processes one level, one row per thread.  Every row performs the
// reference's operations in the reference's order on the same operands, so
// d, d_inv and z = M^-1 r are bit-identical to the serial sweeps.  A
// structured N^3 grid has 3N-2 levels per sweep.
//
// The scaled matrix of the solve is used (a_ij * fn rounded like the
// reference's in-place scaling, Cg.f90:77-84); the stand-alone building
// block passes fn = 1.
#include "tfg_spmv.h"

namespace tfg {

// d_inv_i = a_ii - sum_{k<i} d_k a_ik^2, d_i = 1 / d_inv_i     Prec_Form.f90:68-78
__global__ __launch_bounds__(kBlock) void k_icc_form(IccArgs c, const double* __restrict__ val,
                                                     const State* st, int lo, int hi) {
  const int t = lo + blockIdx.x * kBlock + threadIdx.x;
  if (t >= hi) return;
  const double fn = st ? st->fn : 1.0;
  const int i = c.rows_f[t];
  const int rs = c.row[i], re = c.row[i + 1];
  double sum = val[rs] * fn;
  for (int j = rs + 1; j < re; ++j) {
    const int k = c.col[j];
    if (k > i) break;
    const double a = val[j] * fn;
    sum = sum - c.d[k] * a * a;
  }
  c.dinv[i] = sum;
  c.d[i] = 1.0 / sum;
}

// forward substitution y_i = (r_i - sum_{k<i} a_ik y_k) d_i  Prec_Solve.f90:64-75
__global__ __launch_bounds__(kBlock) void k_icc_fwd(IccArgs c, const double* __restrict__ val,
                                                    const State* st, const double* __restrict__ r,
                                                    int lo, int hi) {
  if (st && st->done) return;
  const int t = lo + blockIdx.x * kBlock + threadIdx.x;
  if (t >= hi) return;
  const double fn = st ? st->fn : 1.0;
  const int i = c.rows_f[t];
  const int rs = c.row[i], re = c.row[i + 1];
  double sum = r[i];
  for (int j = rs + 1; j < re; ++j) {
    const int k = c.col[j];
    if (k > i) break;
    sum = sum - (val[j] * fn) * c.y[k];
  }
  c.y[i] = sum * c.d[i];
}

// z_i = (y_i d_inv_i - sum_{k>i} a_ik z_k) d_i, the upper terms taken from
// the end of the row down                                Prec_Solve.f90:77-92
__global__ __launch_bounds__(kBlock) void k_icc_bwd(IccArgs c, const double* __restrict__ val,
                                                    const State* st, double* __restrict__ z,
                                                    int lo, int hi) {
  if (st && st->done) return;
  const int t = lo + blockIdx.x * kBlock + threadIdx.x;
  if (t >= hi) return;
  const double fn = st ? st->fn : 1.0;
  const int i = c.rows_b[t];
  const int rs = c.row[i], re = c.row[i + 1];
  double sum = c.y[i] * c.dinv[i];
  for (int j = re - 1; j > rs; --j) {
    const int k = c.col[j];
    if (k < i) break;
    if (k < c.n) sum = sum - (val[j] * fn) * z[k];
  }
  z[i] = sum * c.d[i];
}

// CG with ICC, iteration k: the stopping test of iteration k-1 on the update
// partials (one workgroup), before the sweeps (Cg.f90:214-222).
__global__ __launch_bounds__(kBlock) void k_icc_check(CgArgs a, int k) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  State* st = a.st;
  if (st->done) return;
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  gather<5>(a.pt.upd, a.pt.pstride, a.pt.g_upd, ops, s_tmp, s_out);
  if (threadIdx.x != 0) return;
  double res;
  if (st->criterion == TFG_CRIT_RELRES) {
    res = sqrt(s_out[4]) / st->bnorm;
  } else {
    const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(s_out[2] - s_out[1], kTiny);
    res = sqrt(s_out[0]) / range;
  }
  st->res = res;
  st->iter = k - 1;
  if (!isfinite(res)) {
    st->status = TFG_EBREAKDOWN;
    st->niter = k - 1;
    st->done = 1;
  } else if (res < st->tol) {
    st->niter = k - 1;
    st->done = 1;
  }
}

// partials of rho = (z, r)                                  Cg.f90:150-156
__global__ __launch_bounds__(kBlock) void k_icc_rho(CgArgs a, const double* __restrict__ z) {
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  double v[1] = {0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) v[0] = v[0] + z[i] * a.r[i];
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.ext, a.pt.pstride, s_tmp, s_out);
}

// beta = rho_k / rho_{k-1}; p = z + beta p (p = z when k = 1)  Cg.f90:158-171
__global__ __launch_bounds__(kBlock) void k_icc_pupdate(CgArgs a, const double* __restrict__ z,
                                                        int k) {
  __shared__ double s_tmp[4], s_out[1];
  State* st = a.st;
  if (st->done) return;
  const int ops[1] = {OP_SUM};
  gather<1>(a.pt.ext, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  const double rho = s_out[0];
  const double beta = k > 1 ? rho / st->rho[(k - 1) & 1] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rho[k & 1] = rho;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride)
    a.p[i] = k == 1 ? z[i] : z[i] + beta * a.p[i];
}

// z = r * (1 / a_ii) on the unscaled matrix (the building block)
__global__ __launch_bounds__(kBlock) void k_jacobi_apply(int n, const int* __restrict__ row,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ r,
                                                         double* __restrict__ z) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) z[i] = r[i] * (1.0 / val[row[i]]);
}

hipError_t launch_jacobi_apply(hipStream_t s, int n, const int* row, const double* val,
                               const double* r, double* z) {
  k_jacobi_apply<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, row, val, r, z);
  return hipGetLastError();
}

hipError_t launch_icc_form(hipStream_t s, const IccArgs& c, const double* val, const State* st) {
  for (int l = 0; l < c.nlev_f; ++l) {
    const int lo = c.off_f_host[l], hi = c.off_f_host[l + 1];
    k_icc_form<<<(hi - lo + kBlock - 1) / kBlock, kBlock, 0, s>>>(c, val, st, lo, hi);
  }
  return hipGetLastError();
}

hipError_t launch_icc_apply(hipStream_t s, const IccArgs& c, const double* val, const State* st,
                            const double* r, double* z) {
  for (int l = 0; l < c.nlev_f; ++l) {
    const int lo = c.off_f_host[l], hi = c.off_f_host[l + 1];
    k_icc_fwd<<<(hi - lo + kBlock - 1) / kBlock, kBlock, 0, s>>>(c, val, st, r, lo, hi);
  }
  for (int l = 0; l < c.nlev_b; ++l) {
    const int lo = c.off_b_host[l], hi = c.off_b_host[l + 1];
    k_icc_bwd<<<(hi - lo + kBlock - 1) / kBlock, kBlock, 0, s>>>(c, val, st, z, lo, hi);
  }
  return hipGetLastError();
}

hipError_t launch_icc_check(const Launch& L, const CgArgs& a, int k) {
  k_icc_check<<<1, kBlock, 0, L.stream>>>(a, k);
  return hipGetLastError();
}

hipError_t launch_icc_rho(const Launch& L, const CgArgs& a, const double* z) {
  k_icc_rho<<<L.grid_vec, kBlock, 0, L.stream>>>(a, z);
  return hipGetLastError();
}

hipError_t launch_icc_pupdate(const Launch& L, const CgArgs& a, const double* z, int k) {
  k_icc_pupdate<<<L.grid_vec, kBlock, 0, L.stream>>>(a, z, k);
  return hipGetLastError();
}

}  // namespace tfg
