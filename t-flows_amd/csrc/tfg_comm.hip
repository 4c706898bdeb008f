// Multi-rank solves: the matrix of one MPI rank (rows 1..ni owned, rows
// ni+1..nt the buffer cells its rows reference) with the caller's
// communication hooks standing in for Comm_Mod: Global % Sum_Real /
// Min_Real / Max_Real (Shared/Comm_Mod/Parallel/Global/) and
// Grid % Exchange_Cells_Real (Grid_Mod/Parallel/Exchange_Cells_Real.f90:2-129), called where
// Cg.f90 calls them (:72-73, :156, :176, :197, NRMS :38,53-54, :240).
//
// Every global scalar therefore passes through the host once per use: the
// kernels below publish per-workgroup partials as in the one-rank path,
// k_cm_local folds them into this rank's totals (fixed order), the host
// hands those to the caller's allreduce and writes the global results into
// the State for the next kernel.  Products, updates and the row order of
// every sum are the one-rank kernels'.
#include "tfg_spmv.h"

namespace tfg {

// Totals of NV partial rows (fixed order) -> out[NV] (one workgroup).
template <int NV>
__global__ __launch_bounds__(kBlock) void k_cm_local(const double* part, int pstride, int G,
                                                     Ops<NV> ops, double* out) {
  __shared__ double s_tmp[NV * 4], s_out[NV];
  gather<NV>(part, pstride, G, ops.v, s_tmp, s_out);
  if (threadIdx.x < NV) out[threadIdx.x] = s_out[threadIdx.x];
}

// p = z + beta p with the global beta of State (p = z when k = 1)
//                                                    Cg.f90:145, 158-171
// ICC: z is this rank's DIC sweep of r (Prec_Solve.f90:64-92 over rows
// 1..ni, buffer cells left out), passed in zv.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cm_pupdate(CgArgs a, const double* __restrict__ zv,
                                                       int k) {
  const double beta = a.st->beta;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double z = PREC == PREC_ICC ? zv[i] : (PREC == PREC_JACOBI ? a.r[i] * (1.0 / a.ad[i]) : a.r[i]);
    a.p[i] = k == 1 ? z : z + beta * a.p[i];
  }
}

// x += alpha p, r -= alpha q with the global alpha, and this rank's
// partials of the NRMS sum, x range, next (z, r) and (r, r)   Cg.f90:204-218
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_cm_update(CgArgs a) {
  __shared__ double s_tmp[5 * 4], s_out[5];
  const double alpha = a.st->alpha;
  double v[5] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const double d = a.ad[i];
    const double xv = a.x[i] + alpha * a.p[i];
    const double rv = a.r[i] - alpha * a.q[i];
    a.x[i] = xv;
    a.r[i] = rv;
    const double z = PREC == PREC_JACOBI ? rv * (1.0 / d) : rv;
    v[0] = v[0] + (rv * rv) / (d * d);
    v[1] = fmin(v[1], xv);
    v[2] = fmax(v[2], xv);
    v[3] = v[3] + z * rv;
    v[4] = v[4] + rv * rv;
  }
  const int ops[5] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<5>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

hipError_t launch_cm_local(hipStream_t s, const double* part, int pstride, int G, int nv,
                           const int* ops, double* out) {
  switch (nv) {
    case 1: { Ops<1> o{{ops[0]}}; k_cm_local<1><<<1, kBlock, 0, s>>>(part, pstride, G, o, out); break; }
    case 2: { Ops<2> o{{ops[0], ops[1]}}; k_cm_local<2><<<1, kBlock, 0, s>>>(part, pstride, G, o, out); break; }
    case 5: {
      Ops<5> o{{ops[0], ops[1], ops[2], ops[3], ops[4]}};
      k_cm_local<5><<<1, kBlock, 0, s>>>(part, pstride, G, o, out);
      break;
    }
    case 7: {
      Ops<7> o{{ops[0], ops[1], ops[2], ops[3], ops[4], ops[5], ops[6]}};
      k_cm_local<7><<<1, kBlock, 0, s>>>(part, pstride, G, o, out);
      break;
    }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cm_pupdate(const Launch& L, const CgArgs& a, int prec, const double* z, int k) {
  if (prec == PREC_ICC) {
    if (!z) return hipErrorInvalidValue;
    k_cm_pupdate<PREC_ICC><<<L.grid_vec, kBlock, 0, L.stream>>>(a, z, k);
  } else if (prec == PREC_JACOBI) {
    k_cm_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, nullptr, k);
  } else {
    k_cm_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, nullptr, k);
  }
  return hipGetLastError();
}

hipError_t launch_cm_update(const Launch& L, const CgArgs& a, int prec) {
  if (prec == PREC_JACOBI) k_cm_update<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a);
  else k_cm_update<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a);
  return hipGetLastError();
}

}  // namespace tfg
