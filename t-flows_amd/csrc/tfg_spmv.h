// Row engines of the CSR product on gfx950.  Each engine walks the rows a
// workgroup owns and, for every row i, forms
//      acc_i = init(i)  (+|-)  sum_j (a_ij * fn) * x_{c_ij}
// adding the terms ONE AT A TIME IN CSR ORDER, exactly like the reference's
// loop (Cg.f90:177-185, Residual_Vector.f90:34-42).  Compiled without FP
// contraction this reproduces the reference's row sums bit for bit.  The
// engine then calls epi(i, acc_i, x_i, d_i) on exactly one thread, where x_i
// is the x value of the row's first (diagonal) entry and d_i = a_ii * fn.
//
//  StreamEngine<CAP, PIPE> one row per thread; the workgroup first streams
//                         the CSR values/columns of up to 256 consecutive rows
//                         (<= CAP nonzeros) into LDS with coalesced 16-byte
//                         nontemporal loads, then each thread sums its row
//                         from LDS.  With PIPE, while a workgroup sums block b
//                         the CSR streams of its next block are already in
//                         flight into registers (issued after block b's x
//                         gathers, so waiting for those does not wait for them).
//  ScalarStage<CAP>       the same with element-wise staging (unaligned val).
//  WaveEngine<L>          L lanes per row (L = 4, 8, 16), 64/L rows per
//                         wavefront, no LDS and no barriers: lane k of a
//                         group holds term k of a chunk of L terms, and an
//                         ordered scan over the group (DPP row_shr:1, one lane
//                         per step) adds them left to right onto the carried
//                         sum, so the association stays sequential.
// Both deal row blocks to XCDs in contiguous ranges (workgroup g runs on XCD
// g % 8), so each XCD's L2 holds one window of the x gather.
#pragma once

#include "tfg_reduce.h"

namespace tfg {

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef int int4v __attribute__((ext_vector_type(4)));

// Contiguous share [first, end) of `count` items for this workgroup's XCD;
// the XCD's workgroups take every `step`-th item starting at `first`.
struct XcdWalk {
  int first, end, step;
  __device__ __forceinline__ XcdWalk(int count, int per_block = 1) {
    const int G = gridDim.x, g = blockIdx.x;
    if (G >= kXcds && (G % kXcds) == 0) {
      const int x = g % kXcds, l = g / kXcds, nb = G / kXcds;
      first = (int)(((long long)count * x) / kXcds) + l * per_block;
      end = (int)(((long long)count * (x + 1)) / kXcds);
      step = nb * per_block;
    } else {
      first = g * per_block;
      end = count;
      step = G * per_block;
    }
  }
};

// ---------------------------------------------------------------------------
// Stage engines.  The CSR streams of a row block are loaded into registers
// with 16-byte nontemporal loads whose indices are clamped inside the block
// and inside the arrays, so every load is issued unconditionally and back to
// back (a guarded load would make the compiler wait for each one in turn).
// Chunks past the block's end re-read its last chunk (an L1 hit).  The only
// chunks that cannot be read 16 bytes wide are the last 1-3 elements of the
// whole array; those come from uniform scalar loads.  Needs nnz >= 4 and a
// 16-byte aligned val (otherwise ScalarStage is used).
template <int CAP, bool PIPE>
struct StreamEngine {
  static constexpr int NV = (CAP / 2 + 2 + kBlock - 1) / kBlock;  // val chunks per thread
  static constexpr int NC = (CAP / 4 + 2 + kBlock - 1) / kBlock;  // col chunks per thread
  struct Smem {
    double val[CAP + 4];
    int col[CAP + 8];
  };
  struct Regs {
    dbl2 v[NV];     // raw 16-byte chunks (the array tail is patched in put)
    int4v c[NC];
    double tv;      // val[nnz-1]
    int4v tc;       // col[jmax_c+4 ..]
    int r0, nr, s, e, rs, re;
  };

  __device__ __forceinline__ static void fetch(const SpmvShape& m, const double* __restrict__ val,
                                               int nnz, int lb, Regs& g) {
    // Block-uniform facts are read through the constant address space: the
    // arrays are never written during a launch, so these become scalar
    // loads (counted apart from the vector loads in flight).
    typedef __attribute__((address_space(4))) const int cint;
    typedef __attribute__((address_space(4))) const double cdouble;
    const int tid = threadIdx.x;
    const int r0 = ((cint*)m.blk)[lb], r1 = ((cint*)m.blk)[lb + 1];
    const int s = ((cint*)m.row)[r0], e = ((cint*)m.row)[r1];
    g.r0 = r0;
    g.nr = r1 - r0;
    g.s = s;
    g.e = e;
    g.rs = m.row[r0 + min(tid, r1 - r0)];
    g.re = m.row[r0 + min(tid + 1, r1 - r0)];
    // A block that is one long row (e - s > CAP) is summed from global memory
    // and never staged; its loads below still stay inside the arrays.
    const int ov = s & ~1, oc = s & ~3;
    const int jmax_v = (nnz - 2) & ~1, jmax_c = (nnz - 4) & ~3;
    const int lim_v = min((e - 1) & ~1, jmax_v), lim_c = min((e - 1) & ~3, jmax_c);
    // tail elements of the arrays (uniform, scalar loads)
    g.tv = ((cdouble*)val)[nnz - 1];
    g.tc.x = ((cint*)m.col)[min(jmax_c + 4, nnz - 1)];
    g.tc.y = ((cint*)m.col)[min(jmax_c + 5, nnz - 1)];
    g.tc.z = ((cint*)m.col)[min(jmax_c + 6, nnz - 1)];
    g.tc.w = 0;
    // No use of the loaded values here: the loads stay in flight until put().
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int j = ov + 2 * (tid + u * kBlock);
      g.v[u] = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(val + min(j, lim_v)));
    }
#pragma unroll
    for (int u = 0; u < NC; ++u) {
      const int j = oc + 4 * (tid + u * kBlock);
      g.c[u] = __builtin_nontemporal_load(reinterpret_cast<const int4v*>(m.col + min(j, lim_c)));
    }
  }

  // Registers -> LDS, patching the chunk that holds the array's last
  // elements.  Every index written is below e - ov + 2 <= CAP + 3 (val) and
  // e - oc + 4 <= CAP + 7 (col).
  __device__ __forceinline__ static void put(Smem& sm, const Regs& g, int nnz) {
    const int tid = threadIdx.x;
    const int ov = g.s & ~1, oc = g.s & ~3;
    const int nv = (g.e - ov + 1) >> 1, nc = (g.e - oc + 3) >> 2;
    const int jmax_v = (nnz - 2) & ~1, jmax_c = (nnz - 4) & ~3;
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int t = tid + u * kBlock;
      if (t < nv) {
        sm.val[2 * t] = ov + 2 * t > jmax_v ? g.tv : g.v[u].x;
        sm.val[2 * t + 1] = g.v[u].y;
      }
    }
#pragma unroll
    for (int u = 0; u < NC; ++u) {
      const int t = tid + u * kBlock;
      if (t < nc) {
        const int4v c = oc + 4 * t > jmax_c ? g.tc : g.c[u];
        sm.col[4 * t] = c.x;
        sm.col[4 * t + 1] = c.y;
        sm.col[4 * t + 2] = c.z;
        sm.col[4 * t + 3] = c.w;
      }
    }
  }

  template <bool SUB, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem& sm, const SpmvShape& m,
                                             const double* __restrict__ val, double fn,
                                             const double* x, double* s_tmp,
                                             double* s_out, Init init, Epi epi) {
    const int tid = threadIdx.x;
    const int nnz = m.row[m.n];
    // The gather pointer is laundered through an empty asm (keeping the
    // global address space) so its loads are not hoisted or sunk across the
    // prefetch as freely as loads through a __restrict__ pointer.
    typedef __attribute__((address_space(1))) const double gdouble;
    gdouble* xg = (gdouble*)x;
    asm volatile("" : "+s"(xg));
    const XcdWalk w(m.nblk);
    Regs g;
    int lb = w.first;
    if (PIPE && lb < w.end) fetch(m, val, nnz, lb, g);
    while (lb < w.end) {
      const int nxt = lb + w.step;
      if (!PIPE) fetch(m, val, nnz, lb, g);
      const int r0 = g.r0, nr = g.nr, s = g.s, e = g.e, rs = g.rs, re = g.re;
      if (e - s > CAP) {
        // a single row longer than the stage: strided partial sums combined
        // in a fixed order (rows this long do not occur in T-Flows grids)
        double part[1] = {0.0};
        for (int j = s + tid; j < e; j += kBlock) part[0] = part[0] + (val[j] * fn) * x[m.col[j]];
        const int ops[1] = {OP_SUM};
        block_reduce<1>(part, ops, s_tmp, s_out);
        if (tid == 0) {
          const double a0 = init(r0);
          epi(r0, SUB ? a0 - s_out[0] : a0 + s_out[0], x[r0], val[s] * fn);
        }
        __syncthreads();
        if (PIPE) fetch(m, val, nnz, min(nxt, w.end - 1), g);
        lb = nxt;
        continue;
      }
      put(sm, g, nnz);
      __syncthreads();
      const int ov = s & ~1, oc = s & ~3;
      // first eight terms of the row: gathers issued unconditionally with the
      // index clamped into the row (rows are never empty)
      double av[8], xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = min(rs + u, re - 1);
        av[u] = sm.val[j - ov] * fn;
        xv[u] = xg[sm.col[j - oc]];
      }
      // next block's loads, in flight while this block is summed (past the
      // last block the current one is fetched again: no branch, so the
      // compiler can keep them outstanding across the sums).  The scheduling
      // barriers keep them AFTER the gathers: loads complete in issue order,
      // so waiting for a gather must not mean waiting for the prefetch.
      if (PIPE) {
        __builtin_amdgcn_sched_barrier(0);
        fetch(m, val, nnz, min(nxt, w.end - 1), g);
        __builtin_amdgcn_sched_barrier(0);
      }
      // every thread sums (threads past the block's rows on clamped data), so
      // the gathers above have an unconditional use and stay where they are
      const int i = r0 + min(tid, nr - 1);
      double acc = init(i);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (rs + u < re) acc = SUB ? acc - av[u] * xv[u] : acc + av[u] * xv[u];
      if (tid < nr) {
        for (int j = rs + 8; j < re; ++j) {
          const double a = sm.val[j - ov] * fn;
          acc = SUB ? acc - a * x[sm.col[j - oc]] : acc + a * x[sm.col[j - oc]];
        }
        epi(i, acc, xv[0], av[0]);
      }
      __syncthreads();
      lb = nxt;
    }
  }
};

// Stage engine with element-wise staging loads, for a val array that is not
// 16-byte aligned or a matrix with fewer than 4 nonzeros.
template <int CAP>
struct ScalarStage {
  struct Smem {
    double val[CAP];
    int col[CAP];
  };

  template <bool SUB, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem& sm, const SpmvShape& m,
                                             const double* __restrict__ val, double fn,
                                             const double* __restrict__ x, double* s_tmp,
                                             double* s_out, Init init, Epi epi) {
    const int tid = threadIdx.x;
    const XcdWalk w(m.nblk);
    for (int lb = w.first; lb < w.end; lb += w.step) {
      const int r0 = m.blk[lb], nr = m.blk[lb + 1] - r0;
      const int s = m.row[r0], e = m.row[r0 + nr];
      if (e - s > CAP) {
        double part[1] = {0.0};
        for (int j = s + tid; j < e; j += kBlock) part[0] = part[0] + (val[j] * fn) * x[m.col[j]];
        const int ops[1] = {OP_SUM};
        block_reduce<1>(part, ops, s_tmp, s_out);
        if (tid == 0) {
          const double a0 = init(r0);
          epi(r0, SUB ? a0 - s_out[0] : a0 + s_out[0], x[r0], val[s] * fn);
        }
        __syncthreads();
        continue;
      }
      for (int t = tid; t < e - s; t += kBlock) {
        sm.val[t] = val[s + t];
        sm.col[t] = m.col[s + t];
      }
      __syncthreads();
      if (tid < nr) {
        const int i = r0 + tid;
        const int rs = m.row[i], re = m.row[i + 1];
        double acc = init(i);
        for (int j = rs; j < re; ++j) {
          const double a = sm.val[j - s] * fn;
          acc = SUB ? acc - a * x[sm.col[j - s]] : acc + a * x[sm.col[j - s]];
        }
        epi(i, acc, x[sm.col[rs - s]], sm.val[rs - s] * fn);
      }
      __syncthreads();
    }
  }
};

// ---------------------------------------------------------------------------
__device__ __forceinline__ double dpp_shr1(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, false);  // row_shr:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int L>
struct WaveEngine {
  static_assert(L == 4 || L == 8 || L == 16, "lanes per row must divide a DPP row");
  struct Smem {
    int unused;
  };

  template <bool SUB, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem&, const SpmvShape& m,
                                             const double* __restrict__ val, double fn,
                                             const double* __restrict__ x, double*, double*,
                                             Init init, Epi epi) {
    const int jl = threadIdx.x & (L - 1);
    const int gpb = kBlock / L;  // row groups per workgroup
    const XcdWalk w(m.n, gpb);
    for (int i0 = w.first; i0 < w.end; i0 += w.step) {
      const int i = i0 + (int)threadIdx.x / L;
      if (i >= w.end) continue;  // whole groups drop out together
      const int rs = m.row[i], re = m.row[i + 1];
      double acc = jl == 0 ? init(i) : 0.0, xi = 0.0, di = 0.0;
      for (int j0 = rs; j0 < re; j0 += L) {
        const int j = j0 + jl;
        const int jc = min(j, re - 1);  // loads never branch; terms past the end are 0
        const double a = val[jc] * fn, xv = x[m.col[jc]];
        const double t = j < re ? a * xv : 0.0;
        if (j == rs) {
          xi = xv;
          di = a;
        }
        double sacc = SUB ? acc - t : acc + t;  // meaningful on lane 0
#pragma unroll
        for (int k = 1; k < L; ++k) {
          const double prev = dpp_shr1(sacc);
          if (jl == k) sacc = SUB ? prev - t : prev + t;
        }
        // carry = the group's last lane (terms past the row end added +-0)
        acc = __shfl(sacc, ((int)threadIdx.x & 63) | (L - 1), 64);
      }
      if (jl == 0) epi(i, acc, xi, di);
    }
  }
};

}  // namespace tfg
