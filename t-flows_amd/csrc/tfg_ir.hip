// FP32 Krylov sweeps inside FP64 refinement (TFG_FP32_IR), for gfx950.
//
// No counterpart in the reference, which solves in FP64 only.  The CG of
// Cg.f90 runs with FP32 vectors (r, p, q, the correction e) and an FP32
// matrix; alpha, beta and every dot product are FP64 (an FP32 product is
// exact in FP64).  Two things make FP32 work on T-Flows' pressure systems,
// whose pinned Neumann Laplacian has one eigenvalue orders of magnitude
// below the others:
//
//  * the row-excess form of the product: the FP32 matrix holds, in each
//    row's diagonal slot, the row sum s_i = sum_j a_ij (FP64, rounded once)
//    and the product is y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i).  A
//    plain FP32 rounding of the diagonal moves the small eigenvalue by more
//    than its size; this form keeps the operator within FP32 rounding of A
//    in the energy norm, so the near-null space survives;
//  * deflation of the constant vector w (Saad, Yeung, Erhel, Guyomarc'h
//    2000, deflated PCG): with E = w'Aw = sum_i s_i, every direction is
//    made A-orthogonal to w, p = z + beta p - (s'z / E) w, and every restart
//    begins with the coarse correction x += (w'r / E) w, w'r taken as
//    sum_i b_i - (s, x) with compensated row sums s_i and a double-double
//    sum of b.  The FP32 sweeps then never have to resolve the near-null
//    mode, which FP32 rounding keeps re-exciting.  Automatic for symmetric
//    values with E below 1e-2 of the diagonal's sum (TFG_DEFLATE_AUTO).
//
// Refinement (reliable updates with restart): every `check` iterations the
// FP32 residual is looked at; when it has dropped by `delta` since the last
// restart, when its estimate of the stopping measure is below tol, at the
// iteration cap, or every `force` iterations, the correction is folded into
// the FP64 solution, the FP64 true residual b - A x is computed, the
// stopping test is taken on it, and the FP32 CG restarts from it.  All of
// this is decided on the device: the refinement kernels follow every chunk
// and return at once when no refinement is due.  niter counts FP32
// iterations; fin_res is the measure of the last FP64 true residual.
//
// Per FP32 iteration (n rows): pupdate 24n bytes (r, 1/a_ii, p, e), the
// product (symmetric storage: 4(K+1)n + 8n), update 20n (r, q, 1/a_ii, s).
#include <type_traits>

#include "tfg_spmv.h"

namespace tfg {

typedef float flt4 __attribute__((ext_vector_type(4)));

// Partial-sum rows of this form:
//   pt.upd: (z, r), (s, z), (r, r), sum (r / a_ii)^2   -- update and restart
//   pt.ext: sum (r64 / a_ii)^2, x min, x max, (r64, r64), (s, x)  -- true residual
//   pt.irc: sum s_i, sum a_ii fn, sum b_i fn (two rows: double-double) -- conversion
constexpr int kIrUpd = 4;
constexpr int kIrRes = 5;

__device__ __forceinline__ bool ir_lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ SpmvShape ir_shape(const CgArgs& a) {
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

__device__ __forceinline__ double ir_measure(const State* st, double nrms, double xmin,
                                             double xmax, double rr) {
  if (st->criterion == TFG_CRIT_RELRES) return sqrt(rr) / st->bnorm;
  const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(xmax - xmin, kTiny);
  return sqrt(nrms) / range;
}

// ---------------------------------------------------------------------------
// Conversion, once per solve: per row i, s_i = sum_j a_ij fn (the products
// rounded in FP64, summed in CSR order), the FP32 matrix (CSR: fl32(a_ij fn)
// with s_i in the diagonal slot; symmetric storage: s_i in slot 0, the upper
// entries in their slots), s32, rd32 = fl32(1 / (a_ii fn)); partials of
// sum s_i, sum a_ii fn and (when a.b is given: a solve, not the product
// building block) the double-double sum of b_i fn.
template <bool SYM>
__global__ __launch_bounds__(kBlock) void k_ir_convert(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  __shared__ DD s_dd[kBlock / 64 + 1];
  const double fn = a.st->fn;
  double v[2] = {0.0, 0.0};
  DD bs{0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const int rs = a.row[i], re = a.row[i + 1];
    CSum cs{0.0, 0.0};
    for (int j = rs; j < re; ++j) {
      const double aij = a.val[j] * fn;
      cs.add(aij);
      if (j == rs) continue;
      if (SYM) {
        const int d = a.col[j] - i;
#pragma unroll
        for (int k = 0; k < kSymMax; ++k)
          if (k < a.sdg.k && d == a.sdg.off[k]) f.sd32[(long long)(k + 1) * f.stride32 + i] = (float)aij;
      } else {
        f.val32[j] = (float)aij;
      }
    }
    const double s = cs.s + cs.c;
    const double di = a.val[rs] * fn;
    if (SYM) f.sd32[i] = (float)s;
    else f.val32[rs] = (float)s;
    f.s64[i] = s;
    f.s32[i] = (float)s;
    f.rd32[i] = (float)(1.0 / di);
    v[0] = v[0] + s;
    v[1] = v[1] + di;
    if (a.b) bs = dd_add(bs, DD{a.b[i] * fn, 0.0});
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.irc, a.pt.pstride, s_tmp, s_out);
  const DD t = dd_block(bs, s_dd);
  if (threadIdx.x == 0) {
    a.pt.irc[2 * a.pt.pstride + blockIdx.x] = t.hi;
    a.pt.irc[3 * a.pt.pstride + blockIdx.x] = t.lo;
  }
}

// Values symmetric bit for bit (a CSR matrix with the transpose map tpos):
// any entry unlike its mirror clears State::sym_vals (every writer stores
// the same 0).
__global__ __launch_bounds__(kBlock) void k_ir_symcheck(const double* __restrict__ val,
                                                        const int* __restrict__ tpos,
                                                        long long nnz, State* st) {
  bool bad = false;
  for (long long j = blockIdx.x * (long long)kBlock + threadIdx.x; j < nnz;
       j += (long long)gridDim.x * kBlock)
    bad = bad || __double_as_longlong(val[j]) != __double_as_longlong(val[tpos[j]]);
  if (bad) st->sym_vals = 0;
}

// ---------------------------------------------------------------------------
// Iteration k, first phase: beta = rho_k / rho_{k-1} from the partials of
// the previous update (a fresh start right after a restart: beta = 0),
// mu = (s, z) / E (deflation; 0 without), p = z + beta p - mu with
// z = r / a_ii (Jacobi) or r, and the deferred e += alpha_{k-1} p_{k-1} on
// the p read anyway.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_pupdate(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[kIrUpd * 4], s_out[kIrUpd];
  State* st = a.st;
  if (st->done) return;
  const int k = st->kcur;
  const bool fresh = k == st->kref + 1;
  const int stride = gridDim.x * kBlock;
  const int n4 = a.n >> 2;
  flt4* p4 = reinterpret_cast<flt4*>(f.p32);
  flt4* e4 = reinterpret_cast<flt4*>(f.e32);
  const flt4* r4 = reinterpret_cast<const flt4*>(f.r32);
  const flt4* d4 = reinterpret_cast<const flt4*>(f.rd32);
  // this thread's first quadruple, loaded before the sweep of the partials
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  const int tq = n4 > 0 ? min(t0, n4 - 1) : 0;
  flt4 rv0 = {0.f, 0.f, 0.f, 0.f}, dv0 = {1.f, 1.f, 1.f, 1.f}, pv0 = rv0, ev0 = rv0;
  if (n4 > 0) {
    rv0 = r4[tq];
    if (PREC == PREC_JACOBI) dv0 = d4[tq];
    if (!fresh) {
      pv0 = p4[tq];
      ev0 = e4[tq];
    }
  }
  const int ops[kIrUpd] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  gather<kIrUpd>(a.pt.upd, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  const double rho = s_out[0];
  if (!isfinite(rho)) {
    if (ir_lead()) {
      st->status = TFG_EBREAKDOWN;
      st->niter = k - 1;
      st->done = 1;
    }
    return;
  }
  const float beta = fresh ? 0.0f : (float)(rho / st->rho[(k - 1) & 1]);
  const float mu = st->defl_e > 0.0 ? (float)(s_out[1] / st->defl_e) : 0.0f;
  const float ap = fresh ? 0.0f : (float)st->alpha;
  if (ir_lead()) {
    st->rho[k & 1] = rho;
    if (fresh) st->rref = s_out[2];
  }
  auto quad = [&](int t, flt4 rv, flt4 dv, flt4 pv, flt4 ev) {
    const flt4 z = PREC == PREC_JACOBI ? rv * dv : rv;
    if (fresh) {
      p4[t] = z - mu;
    } else {
      e4[t] = ev + ap * pv;
      p4[t] = (z + beta * pv) - mu;
    }
  };
  if (t0 < n4) quad(t0, rv0, dv0, pv0, ev0);
  for (int t = t0 + stride; t < n4; t += stride) {
    const flt4 dv = PREC == PREC_JACOBI ? d4[t] : flt4{1.f, 1.f, 1.f, 1.f};
    quad(t, r4[t], dv, fresh ? flt4{0.f, 0.f, 0.f, 0.f} : p4[t],
         fresh ? flt4{0.f, 0.f, 0.f, 0.f} : e4[t]);
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < (a.n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
    const float z = PREC == PREC_JACOBI ? f.r32[i] * f.rd32[i] : f.r32[i];
    if (fresh) {
      f.p32[i] = z - mu;
    } else {
      f.e32[i] = f.e32[i] + ap * f.p32[i];
      f.p32[i] = (z + beta * f.p32[i]) - mu;
    }
  }
}

// q = A32 p (row-excess form, a CSR engine) and the partials of (p, q).
// The iteration counter moves on here (no workgroup of this kernel reads it).
template <class E>
__global__ __launch_bounds__(kBlock) void k_ir_spmv(CgArgs a, IrArgs f) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  if (ir_lead()) a.st->kcur = a.st->kcur + 1;
  double v[1] = {0.0};
  E::template run<false, true>(
      sm, ir_shape(a), f.val32, 1.0f, f.p32, s_tmp, s_out, [](int) { return 0.0f; },
      [&](int i, float qi, float pi, float) {
        f.q32[i] = qi;
        v[0] = v[0] + (double)pi * (double)qi;
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

// The same with the symmetric FP32 storage (K upper diagonals, MODE as in
// SymEngine32).
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 6)))
void k_ir_spmv_sym(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[4], s_out[1];
  if (a.st->done) return;
  if (ir_lead()) a.st->kcur = a.st->kcur + 1;
  double v[1] = {0.0};
  SymEngine32::run<K, MODE>(
      a.sdg, f.sd32, f.stride32, a.n, f.sorder32, f.nsorder32, f.p32,
      [&](int i, float qi, float pi, bool own) {
        f.q32[i] = qi;
        v[0] = v[0] + (own ? (double)pi * (double)qi : 0.0);
      },
      [&](int i, const float (&q)[4], flt4 p, bool own) {
        flt4 o;
        o.x = q[0];
        o.y = q[1];
        o.z = q[2];
        o.w = q[3];
        *(flt4*)(f.q32 + i) = o;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[0] = v[0] + (own ? (double)p[c] * (double)q[c] : 0.0);
      });
  const int ops[1] = {OP_SUM};
  publish<1>(v, ops, a.pt.spmv, a.pt.pstride, s_tmp, s_out);
}

// alpha = rho_k / (p, q) (FP64); r -= alpha q; partials of the next (z, r),
// (s, z) (deflation), (r, r) and the NRMS sum of the FP32 residual.
template <int PREC, bool DEFL, bool NRMS>
__global__ __launch_bounds__(kBlock) void k_ir_update(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[kIrUpd * 4], s_out[kIrUpd];
  State* st = a.st;
  if (st->done) return;
  const int k = st->kcur - 1;  // the product of this iteration advanced it
  const int stride = gridDim.x * kBlock;
  const int n4 = a.n >> 2;
  flt4* r4 = reinterpret_cast<flt4*>(f.r32);
  const flt4* q4 = reinterpret_cast<const flt4*>(f.q32);
  const flt4* d4 = reinterpret_cast<const flt4*>(f.rd32);
  const flt4* s4 = reinterpret_cast<const flt4*>(f.s32);
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  const int tq = n4 > 0 ? min(t0, n4 - 1) : 0;
  flt4 rv0 = {0.f, 0.f, 0.f, 0.f}, qv0 = rv0, dv0 = rv0, sv0 = rv0;
  if (n4 > 0) {
    rv0 = r4[tq];
    qv0 = q4[tq];
    dv0 = d4[tq];
    if (DEFL) sv0 = s4[tq];
  }
  {
    const int ops1[1] = {OP_SUM};
    gather<1>(a.pt.spmv, a.pt.pstride, a.pt.g_spmv, ops1, s_tmp, s_out);
  }
  const double pq = s_out[0];
  if (pq == 0.0 || !isfinite(pq)) {
    if (ir_lead()) {
      st->status = TFG_EBREAKDOWN;
      st->niter = k;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rho[k & 1] / pq;
  if (ir_lead()) st->alpha = alpha;
  const float a32 = (float)alpha;
  double v[kIrUpd] = {0.0, 0.0, 0.0, 0.0};
  auto one = [&](float& r, float q, float d, float s) {
    r = r - a32 * q;
    const float z = PREC == PREC_JACOBI ? r * d : r;
    v[0] = v[0] + (double)z * (double)r;
    if (DEFL) v[1] = v[1] + (double)s * (double)z;
    v[2] = v[2] + (double)r * (double)r;
    if (NRMS) {
      const double t = (double)(r * d);
      v[3] = v[3] + t * t;
    }
  };
  auto quad = [&](int t, flt4 rv, flt4 qv, flt4 dv, flt4 sv) {
    float r0 = rv.x, r1 = rv.y, r2 = rv.z, r3 = rv.w;
    one(r0, qv.x, dv.x, sv.x);
    one(r1, qv.y, dv.y, sv.y);
    one(r2, qv.z, dv.z, sv.z);
    one(r3, qv.w, dv.w, sv.w);
    flt4 o;
    o.x = r0;
    o.y = r1;
    o.z = r2;
    o.w = r3;
    r4[t] = o;
  };
  if (t0 < n4) quad(t0, rv0, qv0, dv0, sv0);
  for (int t = t0 + stride; t < n4; t += stride)
    quad(t, r4[t], q4[t], d4[t], DEFL ? s4[t] : flt4{0.f, 0.f, 0.f, 0.f});
  if (blockIdx.x == 0 && (int)threadIdx.x < (a.n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
    float r = f.r32[i];
    one(r, f.q32[i], f.rd32[i], DEFL ? f.s32[i] : 0.0f);
    f.r32[i] = r;
  }
  const int ops[kIrUpd] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  publish<kIrUpd>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ---------------------------------------------------------------------------
// ICC in the FP32 sweeps: z = M^-1 r comes from the FP64 DIC sweeps (the
// reference's factor of the scaled matrix, Prec_Form.f90:68-78 /
// Prec_Solve.f90:64-92) on the widened FP32 residual, rounded back to FP32.
// The product and the updates stay FP32; (z, r) and (s, z) are taken here,
// after the sweeps, over rows 0 and 1 of the update's partials (rows 2 and
// 3, (r, r) and the NRMS sum, stay the update's).
__global__ __launch_bounds__(kBlock) void k_ir_widen(CgArgs a, const float* __restrict__ r32,
                                                     double* __restrict__ r64) {
  if (a.st->done) return;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) r64[i] = (double)r32[i];
}

__global__ __launch_bounds__(kBlock) void k_ir_icc_rho(CgArgs a, IrArgs f,
                                                       const double* __restrict__ z64,
                                                       float* __restrict__ z32) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  if (a.st->done) return;
  double v[2] = {0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const float z = (float)z64[i];
    z32[i] = z;
    v[0] = v[0] + (double)z * (double)f.r32[i];
    v[1] = v[1] + (double)f.s32[i] * (double)z;
  }
  const int ops[2] = {OP_SUM, OP_SUM};
  publish<2>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// ---------------------------------------------------------------------------
// Chunk end, first kernel: is a refinement due (State::refine)?  If so,
// x += e + alpha_k p_k in FP64 (the correction with the last iteration's
// deferred step) and e = 0.
__global__ __launch_bounds__(kBlock) void k_ir_fold(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[kIrUpd * 4], s_out[kIrUpd];
  State* st = a.st;
  if (st->done) return;
  const int k = st->kcur - 1;  // the iteration just done
  const int ops[kIrUpd] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  gather<kIrUpd>(a.pt.upd, a.pt.pstride, a.pt.g_vec, ops, s_tmp, s_out);
  const double rr = s_out[2];
  double est;
  if (st->criterion == TFG_CRIT_RELRES) {
    est = sqrt(rr) / st->bnorm;
  } else {
    const double range = st->has_norm ? fmax(st->norm - 0.0, kTiny) : fmax(st->xmax - st->xmin, kTiny);
    est = sqrt(s_out[3]) / range;
  }
  const bool due = rr <= st->delta * st->delta * st->rref || est < st->tol ||
                   (st->force > 0 && k - st->kref >= st->force) || k >= st->miter ||
                   !isfinite(rr);
  if (ir_lead()) st->refine = due ? 1 : 0;
  if (!due) return;
  const float ap = (float)st->alpha;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    a.x[i] = a.x[i] + (double)(f.e32[i] + ap * f.p32[i]);
    f.e32[i] = 0.0f;
  }
}

// Chunk end, second kernel (when a refinement is due, and at the start):
// the FP64 true residual r = b fn - (A fn) x with its sums (the NRMS sum, x
// range, (r, r)) and (s, x) for the deflation's coarse correction.
template <class E>
__global__ __launch_bounds__(kBlock) void k_ir_resid(CgArgs a, const double* __restrict__ s64) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[kIrRes * 4], s_out[kIrRes];
  const State* st = a.st;
  if (st->done || !st->refine) return;
  const double fn = st->fn;
  double v[kIrRes] = {0.0, kHuge, -kHuge, 0.0, 0.0};
  E::template run<true>(
      sm, ir_shape(a), a.val, fn, a.x, s_tmp, s_out, [&](int i) { return a.b[i] * fn; },
      [&](int i, double ri, double xi, double di) {
        a.r[i] = ri;
        v[0] = v[0] + (ri * ri) / (di * di);
        v[1] = fmin(v[1], xi);
        v[2] = fmax(v[2], xi);
        v[3] = v[3] + ri * ri;
        v[4] = v[4] + s64[i] * xi;
      });
  const int ops[kIrRes] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  publish<kIrRes>(v, ops, a.pt.ext, a.pt.pstride, s_tmp, s_out);
}

// Chunk end, third kernel: the stopping test on the true residual (stop:
// State::done with niter = k); otherwise the restart: the deflation's coarse
// correction x += c w with c = w'(b - A x) / E = (sum b - (s, x)) / E (the
// residual's own sum would carry its rows' rounding errors, which 1 / E
// amplifies), r32 = fl32(r - c s), z, and the partials the next
// k_ir_pupdate starts from (its iteration is fresh).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_ir_restart(CgArgs a, IrArgs f) {
  __shared__ double s_tmp[kIrRes * 4], s_out[kIrRes];
  __shared__ DD s_dd[kBlock / 64 + 1];
  State* st = a.st;
  if (st->done || !st->refine) return;
  const int k = st->kcur - 1;
  const int ops5[kIrRes] = {OP_SUM, OP_MIN, OP_MAX, OP_SUM, OP_SUM};
  gather<kIrRes>(a.pt.ext, a.pt.pstride, a.pt.g_spmv, ops5, s_tmp, s_out);
  const double nrms = s_out[0], xmin = s_out[1], xmax = s_out[2], rr = s_out[3], sx = s_out[4];
  const double res = ir_measure(st, nrms, xmin, xmax, rr);
  // the residual of the start (k = 0) is not a refinement
  const int counted = k > 0 ? 1 : 0;
  if (!isfinite(res) || res < st->tol || k >= st->miter) {
    if (ir_lead()) {
      st->nrefine = st->nrefine + counted;
      st->res = res;
      st->iter = k;
      st->niter = k;
      st->status = isfinite(res) ? TFG_OK : TFG_EBREAKDOWN;
      st->done = 1;
    }
    return;
  }
  const int ops2[2] = {OP_SUM, OP_SUM};
  gather<2>(a.pt.irc, a.pt.pstride, a.pt.g_vec, ops2, s_tmp, s_out);
  const double E = s_out[0], sad = s_out[1];
  DD bsum{0.0, 0.0};
  for (int g = threadIdx.x; g < a.pt.g_vec; g += kBlock)
    bsum = dd_add(bsum, DD{a.pt.irc[2 * a.pt.pstride + g], a.pt.irc[3 * a.pt.pstride + g]});
  bsum = dd_block(bsum, s_dd);
  const bool defl = E > 0.0 && (st->defl_mode == TFG_DEFLATE_ON ||
                                (st->defl_mode == TFG_DEFLATE_AUTO && st->sym_vals &&
                                 E < 1e-2 * sad));
  const double c = defl ? ((bsum.hi - sx) + bsum.lo) / E : 0.0;
  if (ir_lead()) {
    st->nrefine = st->nrefine + counted;
    st->res = res;
    st->iter = k;
    st->kref = k;
    st->defl_e = defl ? E : 0.0;
    st->xmin = xmin;
    st->xmax = xmax;
  }
  double v[kIrUpd] = {0.0, 0.0, 0.0, 0.0};
  const bool nrm = st->criterion != TFG_CRIT_RELRES;
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    const float s = f.s32[i], d = f.rd32[i];
    if (defl) a.x[i] = a.x[i] + c;
    const float r = (float)(a.r[i] - c * f.s64[i]);
    f.r32[i] = r;
    f.e32[i] = 0.0f;
    const float z = PREC == PREC_JACOBI ? r * d : r;
    v[0] = v[0] + (double)z * (double)r;
    v[1] = v[1] + (defl ? (double)s * (double)z : 0.0);
    v[2] = v[2] + (double)r * (double)r;
    if (nrm) {
      const double t = (double)(r * d);
      v[3] = v[3] + t * t;
    }
  }
  const int ops[kIrUpd] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  publish<kIrUpd>(v, ops, a.pt.upd, a.pt.pstride, s_tmp, s_out);
}

// x += e after a breakdown (no step pending).
__global__ __launch_bounds__(kBlock) void k_ir_fold_e(CgArgs a, IrArgs f) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
    a.x[i] = a.x[i] + (double)f.e32[i];
    f.e32[i] = 0.0f;
  }
}

// Building block: y = A32 x in the row-excess form.
template <class E>
__global__ __launch_bounds__(kBlock) void k_spmv_lap32(SpmvShape m, const float* __restrict__ val,
                                                       const float* __restrict__ x,
                                                       float* __restrict__ y) {
  __shared__ typename E::Smem sm;
  __shared__ double s_tmp[4], s_out[1];
  E::template run<false, true>(sm, m, val, 1.0f, x, s_tmp, s_out, [](int) { return 0.0f; },
                               [&](int i, float yi, float, float) { y[i] = yi; });
}

template <int K, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 6)))
void k_spmv_lap32_sym(SymDiag g, int n, IrArgs f, const float* __restrict__ x,
                      float* __restrict__ y) {
  SymEngine32::run<K, MODE>(
      g, f.sd32, f.stride32, n, f.sorder32, f.nsorder32, x,
      [&](int i, float yi, float, bool) { y[i] = yi; },
      [&](int i, const float (&q)[4], flt4, bool) {
        flt4 o;
        o.x = q[0];
        o.y = q[1];
        o.z = q[2];
        o.w = q[3];
        *(flt4*)(y + i) = o;
      });
}

// ---------------------------------------------------------------------------
// The FP32 product's engines: the symmetric storage (kind 8) or a CSR
// stream / pattern engine on val32 (the wave engines are not offered).
#define TFG_IR_ENGINE_SWITCH(L, CALL)                                            \
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

template <typename F>
static hipError_t sym32_dispatch(int K, int mode, F f) {
  auto go = [&](auto kc) {
    constexpr int KK = decltype(kc)::value;
    if (mode == 1) f(kc, std::integral_constant<int, 1>{});
    else f(kc, std::integral_constant<int, 0>{});
    (void)KK;
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 5: go(std::integral_constant<int, 5>{}); break;
    case 6: go(std::integral_constant<int, 6>{}); break;
    case 7: go(std::integral_constant<int, 7>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ir_convert(const Launch& L, const Launch& L32, const CgArgs& a,
                             const IrArgs& f, int nnz) {
  (void)nnz;
  if (L32.kind == K_SYM_DIA) k_ir_convert<true><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  else k_ir_convert<false><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_ir_symcheck(hipStream_t s, int grid, const double* val, const int* tpos,
                              long long nnz, State* st) {
  k_ir_symcheck<<<grid, kBlock, 0, s>>>(val, tpos, nnz, st);
  return hipGetLastError();
}

hipError_t launch_ir_widen(const Launch& L, const CgArgs& a, const float* r32, double* r64) {
  k_ir_widen<<<L.grid_vec, kBlock, 0, L.stream>>>(a, r32, r64);
  return hipGetLastError();
}

hipError_t launch_ir_icc_rho(const Launch& L, const CgArgs& a, const IrArgs& f, const double* z64,
                             float* z32) {
  // the grid of the update: every partial of rows 0 and 1 is rewritten
  k_ir_icc_rho<<<a.pt.g_vec, kBlock, 0, L.stream>>>(a, f, z64, z32);
  return hipGetLastError();
}

hipError_t launch_ir_pupdate(const Launch& L, const CgArgs& a, const IrArgs& f) {
  if (f.prec == PREC_JACOBI) k_ir_pupdate<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  else k_ir_pupdate<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_ir_spmv(const Launch& L32, const CgArgs& a, const IrArgs& f) {
  if (L32.kind == K_SYM_DIA) {
    return sym32_dispatch(a.sdg.k, f.quad, [&](auto kc, auto mc) {
      constexpr int K = decltype(kc)::value, MODE = decltype(mc)::value;
      if constexpr (MODE == 1 && K > kSymPairMax) return;
      else k_ir_spmv_sym<K, MODE><<<L32.grid_spmv, kBlock, 0, L32.stream>>>(a, f);
    });
  }
  TFG_IR_ENGINE_SWITCH(L32, (k_ir_spmv<E><<<L32.grid_spmv, kBlock, 0, L32.stream>>>(a, f)));
  return hipGetLastError();
}

hipError_t launch_ir_update(const Launch& L, const CgArgs& a, const IrArgs& f, bool defl,
                            bool nrms) {
  auto go = [&](auto pc, auto dc, auto nc) {
    constexpr int P = decltype(pc)::value;
    constexpr bool D = decltype(dc)::value, N = decltype(nc)::value;
    k_ir_update<P, D, N><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  };
  auto go2 = [&](auto pc) {
    if (defl) {
      if (nrms) go(pc, std::true_type{}, std::true_type{});
      else go(pc, std::true_type{}, std::false_type{});
    } else {
      if (nrms) go(pc, std::false_type{}, std::true_type{});
      else go(pc, std::false_type{}, std::false_type{});
    }
  };
  if (f.prec == PREC_JACOBI) go2(std::integral_constant<int, PREC_JACOBI>{});
  else go2(std::integral_constant<int, PREC_NONE>{});
  return hipGetLastError();
}

hipError_t launch_ir_refine(const Launch& L, const CgArgs& a, const IrArgs& f) {
  k_ir_fold<<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_ir_restart(L, a, f);
}

hipError_t launch_ir_resid(const Launch& L, const CgArgs& a, const IrArgs& f) {
  TFG_ENGINE_SWITCH_T(L, double, (k_ir_resid<E><<<L.grid_spmv, kBlock, 0, L.stream>>>(a, f.s64)));
  return hipGetLastError();
}

hipError_t launch_ir_restart(const Launch& L, const CgArgs& a, const IrArgs& f) {
  hipError_t e = launch_ir_resid(L, a, f);
  if (e != hipSuccess) return e;
  if (f.prec == PREC_JACOBI) k_ir_restart<PREC_JACOBI><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  else k_ir_restart<PREC_NONE><<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_ir_fold(const Launch& L, const CgArgs& a, const IrArgs& f) {
  k_ir_fold_e<<<L.grid_vec, kBlock, 0, L.stream>>>(a, f);
  return hipGetLastError();
}

hipError_t launch_spmv_lap32(const Launch& L32, const SpmvShape& m, const IrArgs& f,
                             const float* x, float* y) {
  if (L32.kind == K_SYM_DIA) {
    return sym32_dispatch(m.sdg.k, f.quad, [&](auto kc, auto mc) {
      constexpr int K = decltype(kc)::value, MODE = decltype(mc)::value;
      if constexpr (MODE == 1 && K > kSymPairMax) return;
      else k_spmv_lap32_sym<K, MODE><<<L32.grid_spmv, kBlock, 0, L32.stream>>>(m.sdg, m.n, f, x, y);
    });
  }
  TFG_IR_ENGINE_SWITCH(L32, (k_spmv_lap32<E><<<L32.grid_spmv, kBlock, 0, L32.stream>>>(m, f.val32, x, y)));
  return hipGetLastError();
}

}  // namespace tfg
