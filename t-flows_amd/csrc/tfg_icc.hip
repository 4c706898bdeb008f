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

// ---------------------------------------------------------------------------
// Box grids: tiled sweeps
// ---------------------------------------------------------------------------
// On a box grid (rows (x, y, z), x slowest; every entry a box neighbour) row
// i's lower entries are, in CSR order, (x-1), (y-1), (z-1), and its upper
// entries from the end of the row down (x+1), (y+1), (z+1).  In sweep
// coordinates (u, v, w) = (x, y, z) for the form and the forward sweep and
// (nx-1-x, ny-1-y, nz-1-z) for the backward sweep, a row reads the finished
// values at u-1, v-1, w-1, in that order.  The grid is cut into tiles of at
// most kBoxTX x kBoxTY x kBoxTZ cells; a tile needs the tiles before it along u, v and w, so
// the tiles of one level X + Y + Z run in one launch, one workgroup each
// (61 launches per sweep at 200^3 instead of 598 row levels).  A workgroup
// loads the three finished faces next to its tile and its cells' values
// into LDS, one thread per cell; one wave then runs the tile's own levels
// lu + lv + lw, a lane per (u, v) pencil, with no barrier between them.
// Every row does the reference's operations in the reference's order; an
// entry the row lacks is a zero in the diagonals and takes a zero product
// away, so for finite values the results are the same bits as the
// row-level sweeps'.

// the seven diagonals of val (entries a row lacks stay zero from the
// allocation)
__global__ __launch_bounds__(kBlock) void k_box_pack(IccArgs c, const double* __restrict__ val) {
  const BoxIcc& b = c.bx;
  const int nyz = b.ny * b.nz;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < c.n; i += stride) {
    const int rs = c.row[i], re = c.row[i + 1];
    b.dg[i] = val[rs];
    for (int j = rs + 1; j < re; ++j) {
      const int off = c.col[j] - i;
      const double v = val[j];
      if (off == -nyz) b.lo[0][i] = v;
      else if (off == -b.nz) b.lo[1][i] = v;
      else if (off == -1) b.lo[2][i] = v;
      else if (off == 1) b.up[2][i] = v;
      else if (off == b.nz) b.up[1][i] = v;
      else if (off == nyz) b.up[0][i] = v;
    }
  }
}

// MODE 0 form: d_inv, d from a_ii and the lower entries   Prec_Form.f90:68-78
// MODE 1 forward: y from r                                Prec_Solve.f90:64-75
// MODE 2 backward: z from y                               Prec_Solve.f90:77-92
// The value the neighbours read (d, y, z) is `out`.  One tile per call of
// the pieces below, by a workgroup of kBoxThreads threads: its cells'
// values into LDS, the three finished faces next to it, the tile's levels
// on one wave, the results back to HBM.
template <int MODE>
struct BoxTile {
  static constexpr int TX = kBoxTX, TY = kBoxTY, TZ = kBoxTZ;
  static_assert(TX * TY == 64, "one lane of the level wave per (u, v) pencil");
  static_assert(TX * TY * TZ == kBoxThreads, "one thread per cell when staging");
  static_assert(TX * TZ >= TY * TZ && TX * TZ >= TX * TY, "the largest face is the v face");
  // LDS strides of the tile with its faces and the pencil stride of the
  // staged cell data, padded so that the lanes of one level fall on
  // different banks
  static constexpr int SW = 1, SV = TZ + 2, SU = (TY + 1) * SV + 2, P = TZ + 2;
  struct Smem {
    double s[(TX + 1) * SU];
    double cv[5][TX * TY * P];  // c0, c1, c2, a0 (r or y), d; MODE 0: d_inv in slot 4
  };
  int u0, v0, w0;
  int su, sv, sw, sq, si;
  bool mine;

  __device__ __forceinline__ BoxTile(const BoxIcc& b, int tile) {
    const int X = tile / (b.nty * b.ntz), Y = (tile / b.ntz) % b.nty, Z = tile % b.ntz;
    u0 = X * b.tx;
    v0 = Y * b.ty;
    w0 = Z * b.tz;
    const int tid = threadIdx.x;
    su = tid / (TY * TZ);
    sv = (tid / TZ) % TY;
    sw = tid % TZ;
    mine = in_tile(b, su, sv, sw);
    sq = (su * TY + sv) * P + sw;
    si = mine ? row_of(b, u0 + su, v0 + sv, w0 + sw) : 0;
  }
  __device__ __forceinline__ static int row_of(const BoxIcc& b, int u, int v, int w) {
    const int x = MODE == 2 ? b.nx - 1 - u : u;
    const int y = MODE == 2 ? b.ny - 1 - v : v;
    const int z = MODE == 2 ? b.nz - 1 - w : w;
    return (x * b.ny + y) * b.nz + z;
  }
  __device__ __forceinline__ bool in_tile(const BoxIcc& b, int lu, int lv, int lw) const {
    return lu < b.tx && lv < b.ty && lw < b.tz && u0 + lu < b.nx && v0 + lv < b.ny &&
           w0 + lw < b.nz;
  }
  // this thread's cell (the tile's natural order, w fastest: the loads
  // coalesce); inputs written by earlier kernels only
  __device__ __forceinline__ void stage_cells(const IccArgs& c, const double* __restrict__ in,
                                              Smem& m) const {
    const BoxIcc& b = c.bx;
    const double* const* cf = MODE == 2 ? b.up : b.lo;
    if (!mine) return;
    m.cv[0][sq] = cf[0][si];
    m.cv[1][sq] = cf[1][si];
    m.cv[2][sq] = cf[2][si];
    if (MODE == 0) {
      m.cv[3][sq] = b.dg[si];
    } else if (MODE == 1) {
      m.cv[3][sq] = in[si];
      m.cv[4][sq] = c.d[si];
    } else {
      m.cv[3][sq] = in[si] * c.dinv[si];
      m.cv[4][sq] = c.d[si];
    }
  }
  // finished faces at u0-1, v0-1, w0-1 (zero outside the grid)
  __device__ __forceinline__ void stage_faces(const BoxIcc& b, const double* out, Smem& m) const {
    const int tid = threadIdx.x;
    for (int t = tid; t < b.ty * b.tz; t += kBoxThreads) {
      const int lv = t / b.tz, lw = t % b.tz, v = v0 + lv, w = w0 + lw;
      m.s[(lv + 1) * SV + (lw + 1) * SW] =
          (u0 > 0 && v < b.ny && w < b.nz) ? out[row_of(b, u0 - 1, v, w)] : 0.0;
    }
    for (int t = tid; t < b.tx * b.tz; t += kBoxThreads) {
      const int lu = t / b.tz, lw = t % b.tz, u = u0 + lu, w = w0 + lw;
      m.s[(lu + 1) * SU + (lw + 1) * SW] =
          (v0 > 0 && u < b.nx && w < b.nz) ? out[row_of(b, u, v0 - 1, w)] : 0.0;
    }
    for (int t = tid; t < b.tx * b.ty; t += kBoxThreads) {
      const int lu = t / b.ty, lv = t % b.ty, u = u0 + lu, v = v0 + lv;
      m.s[(lu + 1) * SU + (lv + 1) * SV] =
          (w0 > 0 && u < b.nx && v < b.ny) ? out[row_of(b, u, v, w0 - 1)] : 0.0;
    }
  }
  // one wave runs the levels: at level l the lane of pencil (pu, pv)
  // computes cell w = l - pu - pv, so a level needs no barrier
  __device__ __forceinline__ void levels(const BoxIcc& b, double fn, Smem& m) const {
    if (threadIdx.x >= TX * TY) return;
    const int pu = threadIdx.x / TY, pv = threadIdx.x % TY;
    const int nlev = b.tx + b.ty + b.tz - 2;
    for (int l = 0; l < nlev; ++l) {
      const int lw = l - pu - pv;
      if (lw >= 0 && in_tile(b, pu, pv, lw)) {
        const int q = (pu * TY + pv) * P + lw;
        const int p = (pu + 1) * SU + (pv + 1) * SV + (lw + 1) * SW;
        const double n0 = m.s[p - SU], n1 = m.s[p - SV], n2 = m.s[p - SW];
        double r;
        if (MODE == 0) {
          double sum = m.cv[3][q] * fn;
          double a = m.cv[0][q] * fn;
          sum = sum - n0 * a * a;
          a = m.cv[1][q] * fn;
          sum = sum - n1 * a * a;
          a = m.cv[2][q] * fn;
          sum = sum - n2 * a * a;
          m.cv[4][q] = sum;
          r = 1.0 / sum;
        } else {
          double sum = m.cv[3][q];
          sum = sum - (m.cv[0][q] * fn) * n0;
          sum = sum - (m.cv[1][q] * fn) * n1;
          sum = sum - (m.cv[2][q] * fn) * n2;
          r = sum * m.cv[4][q];
        }
        m.s[p] = r;
      }
      // the next level reads cells other lanes of this wave stored: release
      // this level's LDS stores and acquire the others' at wavefront scope
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __device__ __forceinline__ void store(const IccArgs& c, double* out, const Smem& m) const {
    if (!mine) return;
    out[si] = m.s[(su + 1) * SU + (sv + 1) * SV + (sw + 1) * SW];
    if (MODE == 0) c.dinv[si] = m.cv[4][sq];
  }
};

// One launch per tile level: the tiles of level toff[l] .. are
// blockIdx.x + lo.
template <int MODE>
__global__ __launch_bounds__(kBoxThreads) void k_box_sweep(IccArgs c, const State* st,
                                                           const double* __restrict__ in,
                                                           double* __restrict__ out, int lo) {
  using T = BoxTile<MODE>;
  __shared__ typename T::Smem m;
  if (MODE != 0 && st && st->done) return;
  const BoxIcc& b = c.bx;
  const double fn = st ? st->fn : 1.0;
  const T t(b, b.tiles[lo + blockIdx.x]);
  t.stage_faces(b, out, m);
  t.stage_cells(c, in, m);
  __syncthreads();
  t.levels(b, fn, m);
  __syncthreads();
  t.store(c, out, m);
}

hipError_t launch_box_sweep(hipStream_t s, const IccArgs& c, int mode, const State* st,
                            const double* in, double* out) {
  // the level loop hands cells between the lanes of one wave of 64
  static const int wave = [] {
    int dev = 0, w = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&w, hipDeviceAttributeWarpSize, dev) != hipSuccess)
      return 0;
    return w;
  }();
  if (wave != 64) return hipErrorInvalidDeviceFunction;
  const int nl = c.bx.ntx + c.bx.nty + c.bx.ntz - 2;
  for (int l = 0; l < nl; ++l) {
    const int lo = c.toff_host[l], cnt = c.toff_host[l + 1] - lo;
    if (mode == 0) k_box_sweep<0><<<cnt, kBoxThreads, 0, s>>>(c, st, in, out, lo);
    else if (mode == 1) k_box_sweep<1><<<cnt, kBoxThreads, 0, s>>>(c, st, in, out, lo);
    else k_box_sweep<2><<<cnt, kBoxThreads, 0, s>>>(c, st, in, out, lo);
  }
  return hipGetLastError();
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
  if (c.box) {
    const int grid = std::max(1, std::min((c.n + kBlock - 1) / kBlock, 4096));
    k_box_pack<<<grid, kBlock, 0, s>>>(c, val);
    return launch_box_sweep(s, c, 0, st, nullptr, c.d);
  }
  for (int l = 0; l < c.nlev_f; ++l) {
    const int lo = c.off_f_host[l], hi = c.off_f_host[l + 1];
    k_icc_form<<<(hi - lo + kBlock - 1) / kBlock, kBlock, 0, s>>>(c, val, st, lo, hi);
  }
  return hipGetLastError();
}

hipError_t launch_icc_apply(hipStream_t s, const IccArgs& c, const double* val, const State* st,
                            const double* r, double* z) {
  if (c.box) {  // the diagonals were packed from val by launch_icc_form
    const hipError_t e = launch_box_sweep(s, c, 1, st, r, c.y);
    if (e != hipSuccess) return e;
    return launch_box_sweep(s, c, 2, st, c.y, z);
  }
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
