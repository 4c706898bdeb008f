// Row engines of the CSR product on gfx950.  Each engine walks the rows a
// workgroup owns and, for every row i, forms
//      acc_i = init(i)  (+|-)  sum_j (a_ij * fn) * x_{c_ij}
// adding the terms ONE AT A TIME IN CSR ORDER, exactly like the reference's
// loop (Cg.f90:177-185, Residual_Vector.f90:34-42).  Compiled without FP
// contraction this reproduces the reference's row sums bit for bit.  The
// engine then calls epi(i, acc_i, x_i, d_i) on exactly one thread, where x_i
// is the x value of the row's first (diagonal) entry and d_i = a_ii * fn.
// T is the value type of the matrix and of x (double, or float for the
// inner solves of iterative refinement).
//
//  StreamEngine<CAP, PIPE, T>  one row per thread; the workgroup first
//                         streams the CSR values/columns of up to 256
//                         consecutive rows (<= CAP nonzeros) into LDS with
//                         coalesced 16-byte nontemporal loads, then each thread
//                         sums its row from LDS.  With PIPE, while a workgroup
//                         sums block b the CSR streams of its next block are
//                         already in flight into registers (issued after block
//                         b's x gathers, so waiting for those does not wait
//                         for them: vector loads complete in issue order).
//  ScalarStage<CAP, T>    the same with element-wise staging (unaligned val).
//  WaveEngine<L, T>       L lanes per row (L = 4, 8, 16), 64/L rows per
//                         wavefront, no LDS and no barriers: lane k of a
//                         group holds term k of a chunk of L terms, and an
//                         ordered scan over the group (DPP row_shr:1, one lane
//                         per step) adds them left to right onto the carried
//                         sum, so the association stays sequential.
// All deal row blocks to XCDs in contiguous ranges (workgroup g runs on XCD
// g % 8), so each XCD's L2 holds one window of the x gather.
#pragma once

#include <type_traits>

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
// A 16-byte-chunk stream of one CSR array (val or col) for one row block
// [s, e): chunk t of the workgroup covers elements o + EPC*t .. + EPC-1 with
// o = s rounded down to a chunk.  Loads are issued unconditionally with the
// chunk index clamped inside the block and inside the array, so the compiler
// can keep them all in flight (a guarded load makes it wait for each in
// turn); chunks past the block re-read its last chunk (an L1 hit).  The one
// chunk that holds the array's last 1..EPC-1 elements cannot be read 16 bytes
// wide; those elements come from uniform scalar loads (`tail`) and are
// patched in when the chunk is written to LDS.  Needs nnz >= EPC.
template <typename E, int NCH>
struct ChunkStream {
  static constexpr int EPC = 16 / (int)sizeof(E);
  typedef E V __attribute__((ext_vector_type(16 / sizeof(E))));
  V v[NCH];
  V tail;

  __device__ __forceinline__ void fetch(const E* __restrict__ arr, int nnz, int s, int e) {
    typedef __attribute__((address_space(4))) const E cE;
    const int tid = threadIdx.x;
    const int o = s & ~(EPC - 1);
    const int jmax = (nnz - EPC) & ~(EPC - 1);
    const int lim = min((e - 1) & ~(EPC - 1), jmax);
#pragma unroll
    for (int k = 0; k < EPC; ++k) tail[k] = ((cE*)arr)[min(jmax + EPC + k, nnz - 1)];
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int j = o + EPC * (tid + u * kBlock);
      v[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(arr + min(j, lim)));
    }
  }

  // Registers -> LDS.  Writes indices below e - o + EPC - 1 <= CAP + 2*EPC - 2.
  __device__ __forceinline__ void put(E* lds, int nnz, int s, int e) const {
    const int tid = threadIdx.x;
    const int o = s & ~(EPC - 1);
    const int n = (e - o + EPC - 1) / EPC;
    const int jmax = (nnz - EPC) & ~(EPC - 1);
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int t = tid + u * kBlock;
      if (t < n) {
        const V c = o + EPC * t > jmax ? tail : v[u];
        if constexpr (sizeof(E) == 1) {
          reinterpret_cast<V*>(lds)[t] = c;  // one 16-byte LDS store
        } else {
#pragma unroll
          for (int k = 0; k < EPC; ++k) lds[EPC * t + k] = c[k];
        }
      }
    }
  }
};

// Sum of a row longer than the stage: strided partial sums combined in a
// fixed order (rows this long do not occur in T-Flows grids).
template <bool SUB, bool LAP = false, typename T, typename Init, typename Epi>
__device__ __forceinline__ void long_row(const SpmvShape& m, const T* __restrict__ val, T fn,
                                         const T* __restrict__ x, int r0, int s, int e,
                                         double* s_tmp, double* s_out, Init& init, Epi& epi) {
  double part[1] = {0.0};
  for (int j = s + (int)threadIdx.x; j < e; j += kBlock)
    part[0] = part[0] + (double)((val[j] * fn) * (LAP && j != s ? x[m.col[j]] - x[r0]
                                                                 : x[m.col[j]]));
  const int ops[1] = {OP_SUM};
  block_reduce<1>(part, ops, s_tmp, s_out);
  if (threadIdx.x == 0) {
    const T a0 = init(r0);
    const T sum = (T)s_out[0];
    epi(r0, SUB ? a0 - sum : a0 + sum, x[r0], val[s] * fn);
  }
  __syncthreads();
}

// Term of a row sum: a x_j, or in the row-excess form (LAP, the FP32 sweeps
// of iterative refinement) a (x_j - x_i) for every entry after the first,
// the diagonal, whose value is then the row's excess s_i = sum_j a_ij and
// takes x_i: s_i x_i + sum_{j != i} a_ij (x_j - x_i) = sum_j a_ij x_j.
template <bool LAP, typename T>
__device__ __forceinline__ T row_term(T a, T xj, T xi, bool first) {
  return (LAP && !first) ? a * (xj - xi) : a * xj;
}

// ---------------------------------------------------------------------------
// CODED: the column stream is the 1-byte code of each nonzero's offset
// col - row in a per-matrix table of at most 256 distinct offsets (a
// structured grid has 7), a quarter of the int32 column traffic.
template <int CAP, bool PIPE, typename T, bool CODED = false>
struct StreamEngine {
  typedef typename std::conditional<CODED, uint8_t, int>::type C;
  static constexpr int EPV = 16 / (int)sizeof(T);
  static constexpr int EPC = 16 / (int)sizeof(C);
  static constexpr int NV = (CAP / EPV + 2 + kBlock - 1) / kBlock;  // val chunks per thread
  static constexpr int NC = (CAP / EPC + 2 + kBlock - 1) / kBlock;  // col chunks per thread
  struct Smem {
    T val[CAP + 2 * EPV];
    __attribute__((aligned(16))) C col[CAP + 2 * EPC];
    int tab[CODED ? 256 : 1];
  };
  struct Regs {
    ChunkStream<T, NV> v;
    ChunkStream<C, NC> c;
    int r0, nr, s, e, rs, re;
  };

  __device__ __forceinline__ static void fetch(const SpmvShape& m, const T* __restrict__ val,
                                               int nnz, int lb, Regs& g) {
    // Block-uniform facts are read through the constant address space: the
    // arrays are never written during a launch, so these become scalar
    // loads (counted apart from the vector loads in flight).
    typedef __attribute__((address_space(4))) const int cint;
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
    // and never staged; its loads here still stay inside the arrays.  No use
    // of the loaded values here: they stay in flight until put().
    g.v.fetch(val, nnz, s, e);
    if constexpr (CODED)
      g.c.fetch(m.code, nnz, s, e);
    else
      g.c.fetch(m.col, nnz, s, e);
  }

  // column of staged entry j of row i
  __device__ __forceinline__ static int column(const Smem& sm, int i, int j, int oc) {
    if constexpr (CODED)
      return i + sm.tab[sm.col[j - oc]];
    else
      return sm.col[j - oc];
  }

  template <bool SUB, bool LAP = false, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem& sm, const SpmvShape& m,
                                             const T* __restrict__ val, T fn, const T* x,
                                             double* s_tmp, double* s_out, Init init, Epi epi) {
    const int tid = threadIdx.x;
    const int nnz = m.row[m.n];
    if constexpr (CODED) {
      for (int t = tid; t < m.ntab; t += kBlock) sm.tab[t] = m.tab[t];
      __syncthreads();
    }
    // The gather pointer is laundered through an empty asm (keeping the
    // global address space) so its loads are not hoisted or sunk across the
    // prefetch as freely as loads through a __restrict__ pointer.
    typedef __attribute__((address_space(1))) const T gT;
    gT* xg = (gT*)x;
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
        long_row<SUB, LAP>(m, val, fn, x, r0, s, e, s_tmp, s_out, init, epi);
        if (PIPE) fetch(m, val, nnz, min(nxt, w.end - 1), g);
        lb = nxt;
        continue;
      }
      g.v.put(sm.val, nnz, s, e);
      g.c.put(sm.col, nnz, s, e);
      __syncthreads();
      const int ov = s & ~(EPV - 1), oc = s & ~(EPC - 1);
      // the thread's row (threads past the block's rows take its last row)
      const int i = r0 + min(tid, nr - 1);
      // first eight terms of the row: gathers issued unconditionally with the
      // index clamped into the row (rows are never empty)
      T av[8], xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = min(rs + u, re - 1);
        av[u] = sm.val[j - ov] * fn;
        xv[u] = xg[column(sm, i, j, oc)];
      }
      // next block's loads, in flight while this block is summed (past the
      // last block the current one is fetched again: no branch, so the
      // compiler can keep them outstanding across the sums).  The scheduling
      // barriers keep them AFTER the gathers.
      if (PIPE) {
        __builtin_amdgcn_sched_barrier(0);
        fetch(m, val, nnz, min(nxt, w.end - 1), g);
        __builtin_amdgcn_sched_barrier(0);
      }
      // every thread sums (threads past the block's rows on clamped data), so
      // the gathers above have an unconditional use and stay where they are
      T acc = init(i);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (rs + u < re) {
          const T t = row_term<LAP>(av[u], xv[u], xv[0], u == 0);
          acc = SUB ? acc - t : acc + t;
        }
      if (tid < nr) {
        for (int j = rs + 8; j < re; ++j) {
          const T a = sm.val[j - ov] * fn;
          const T t = row_term<LAP>(a, x[column(sm, i, j, oc)], xv[0], false);
          acc = SUB ? acc - t : acc + t;
        }
        epi(i, acc, xv[0], av[0]);
      }
      __syncthreads();
      lb = nxt;
    }
  }
};

// Exclusive prefix sum of v over the workgroup (thread order); s_wave needs
// kBlock/64 ints.  Contains barriers: every thread must call it.
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wid; ++w) base += s_wave[w];
  __syncthreads();
  return base + inc - v;
}

// Rows described by a pattern: the column offsets (col - row, in CSR order)
// of a row are entry pat[row] of a per-matrix table of at most kPatCount
// offset sequences of at most kPatMax entries (a 7-point box grid has 27:
// one per boundary combination).  So neither a column stream nor the row
// pointers are read, only the values and one byte per row; a row's first
// nonzero is the block's first plus the prefix sum of the pattern lengths
// before it (one scan per block).  Otherwise the pipelined stream engine.
template <int CAP, bool PIPE, typename T>
struct PatternEngine {
  static constexpr int EPV = 16 / (int)sizeof(T);
  static constexpr int NV = (CAP / EPV + 2 + kBlock - 1) / kBlock;
  struct Smem {
    T val[CAP + 2 * EPV];
    int off[kPatCount * kPatMax];
    int len[kPatCount];
    int wave[kBlock / 64];
  };
  struct Regs {
    ChunkStream<T, NV> v;
    int r0, nr, s, e, pat;
  };

  __device__ __forceinline__ static void fetch(const SpmvShape& m, const T* __restrict__ val,
                                               int nnz, int lb, Regs& g) {
    typedef __attribute__((address_space(4))) const int cint;
    const int r0 = ((cint*)m.blk)[lb], r1 = ((cint*)m.blk)[lb + 1];
    const int s = ((cint*)m.row)[r0], e = ((cint*)m.row)[r1];
    g.r0 = r0;
    g.nr = r1 - r0;
    g.s = s;
    g.e = e;
    g.pat = m.pat[r0 + min((int)threadIdx.x, r1 - r0 - 1)];
    g.v.fetch(val, nnz, s, e);
  }

  template <bool SUB, bool LAP = false, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem& sm, const SpmvShape& m,
                                             const T* __restrict__ val, T fn, const T* x,
                                             double* s_tmp, double* s_out, Init init, Epi epi) {
    (void)s_tmp;
    (void)s_out;
    const int tid = threadIdx.x;
    const int nnz = m.row[m.n];
    for (int t = tid; t < kPatCount * kPatMax; t += kBlock) sm.off[t] = m.poff[t];
    for (int t = tid; t < kPatCount; t += kBlock) sm.len[t] = m.plen[t];
    __syncthreads();
    typedef __attribute__((address_space(1))) const T gT;
    gT* xg = (gT*)x;
    asm volatile("" : "+s"(xg));
    const XcdWalk w(m.nblk);
    Regs g;
    int lb = w.first;
    if (PIPE && lb < w.end) fetch(m, val, nnz, lb, g);
    while (lb < w.end) {
      const int nxt = lb + w.step;
      if (!PIPE) fetch(m, val, nnz, lb, g);
      const int r0 = g.r0, nr = g.nr, s = g.s, e = g.e, p = g.pat;
      g.v.put(sm.val, nnz, s, e);
      // row starts: prefix sum of the pattern lengths (threads past the
      // block's rows take its last row, whose start is e - L)
      const int L = sm.len[p];
      const int pre = block_exclusive_scan(tid < nr ? L : 0, sm.wave);
      const int rs = tid < nr ? s + pre : e - L;
      const int i = r0 + min(tid, nr - 1);
      const int ov = s & ~(EPV - 1);
      const int* po = sm.off + p * kPatMax;
      T av[8], xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int uu = min(u, L - 1);
        av[u] = sm.val[rs + uu - ov] * fn;
        xv[u] = xg[i + po[uu]];
      }
      if (PIPE) {
        __builtin_amdgcn_sched_barrier(0);
        fetch(m, val, nnz, min(nxt, w.end - 1), g);
        __builtin_amdgcn_sched_barrier(0);
      }
      T acc = init(i);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < L) {
          const T t = row_term<LAP>(av[u], xv[u], xv[0], u == 0);
          acc = SUB ? acc - t : acc + t;
        }
      if (tid < nr) {
        for (int u = 8; u < L; ++u) {
          const T a = sm.val[rs + u - ov] * fn;
          const T t = row_term<LAP>(a, x[i + po[u]], xv[0], false);
          acc = SUB ? acc - t : acc + t;
        }
        epi(i, acc, xv[0], av[0]);
      }
      __syncthreads();
      lb = nxt;
    }
  }
};

// Symmetric matrices whose entries lie on a few diagonals: a(i, j) =
// a(j, i) bit for bit, and every column offset j - i is 0 or +-d_k for one
// set 0 < d_0 < ... < d_{K-1} (K <= kSymMax).  The values are held once, as
// the diagonal and the K upper diagonals, each zero where a row has no such
// entry (SpmvShape::sd): per product 8 B per row and per upper diagonal
// instead of 8 B per nonzero, and no per-row structure at all.  Row i sums
//   a_ii x_i,
//   a(i, i - d_k) x_{i - d_k} for k = K-1 .. 0 (the mirror: slot k of row
//     i - d_k),
//   a(i, i + d_k) x_{i + d_k} for k = 0 .. K-1,
// which is the CSR order (diagonal first, then columns ascending; checked
// when the handle is made), with each value scaled as the stored one is.  A
// missing entry adds a zero product and leaves the sum as it was, so for
// finite x the products are bit-identical to the CSR engines' (the x index
// of a missing entry is clamped into the vector).  Row blocks come in a
// plane-interleaved order built on the host (sorder).
// Modes: 0 one row per thread, blocks of at most 256 rows; 1 two rows per
// thread (d_0 = 1, the other offsets even, n even; blocks of at most 512
// rows at even starts): rows i and i + 1 (i even) read 16-byte pairs, the
// +-1 neighbours from the pairs at i - 2 and i + 2 and every other term
// from one aligned pair; 2 one row per thread over mode 1's blocks (x or y
// not 16-byte aligned).  The epilogue takes a fifth argument: whether the
// lane's row is its own (a lane past the block's rows repeats the last row
// and must not count it twice in a reduction).
struct SymEngine {
  __device__ __forceinline__ static int2 ids(const SpmvShape& m, int lb) {
    typedef __attribute__((address_space(4))) const int cint;
    return make_int2(((cint*)m.sorder)[2 * lb], ((cint*)m.sorder)[2 * lb + 1]);
  }

  template <bool SUB>
  __device__ __forceinline__ static double term(double acc, double v, double fn, double x) {
    const double a = v * fn;
    return SUB ? acc - a * x : acc + a * x;
  }

  // Loads at a workgroup-uniform base plus a 32-bit byte offset (one VGPR
  // per address; sdg.stride < 2^29 is checked on the host).
  typedef __attribute__((address_space(1))) const char gB;
  template <typename T>
  __device__ __forceinline__ static T at(gB* base, unsigned byte) {
    return *(__attribute__((address_space(1))) const T*)(base + byte);
  }
  template <typename T>  // non-temporal: no other row reads it
  __device__ __forceinline__ static T at_nt(gB* base, unsigned byte) {
    return __builtin_nontemporal_load((__attribute__((address_space(1))) const T*)(base + byte));
  }

  // rows r0 + h * kBlock + tid, h < H, of block [r0, r1)
  template <bool SUB, int K, int H, typename Init, typename Epi>
  __device__ __forceinline__ static void rows(const SpmvShape& m, double fn, const double* x,
                                              int r0, int r1, Init& init, Epi& epi) {
    gB* xb = (gB*)x;
    gB* sb = (gB*)m.sd;
    const long long S8 = 8 * m.sdg.stride;
    const unsigned lastb = 8u * (unsigned)(m.n - 1);
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int t = r0 + h * kBlock + (int)threadIdx.x;
      const int i = min(t, r1 - 1);
      const unsigned ib = 8u * (unsigned)i;
      double lo[K], up[K], xl[K], xu[K];
      const double d = at_nt<double>(sb, ib);
      const double x0 = at<double>(xb, ib);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const unsigned ob = 8u * (unsigned)m.sdg.off[k];
        gB* slot = sb + (k + 1) * S8;
        up[k] = at<double>(slot, ib);
        lo[k] = at<double>(slot - ob, ib);  // in the zeros after slot k - 1 for i < d_k
        xu[k] = at<double>(xb, min(ib + ob, lastb));
        xl[k] = at<double>(xb, max(ib, ob) - ob);
      }
      double acc = init(i);
      acc = term<SUB>(acc, d, fn, x0);
#pragma unroll
      for (int k = K - 1; k >= 0; --k) acc = term<SUB>(acc, lo[k], fn, xl[k]);
#pragma unroll
      for (int k = 0; k < K; ++k) acc = term<SUB>(acc, up[k], fn, xu[k]);
      epi(i, acc, x0, d * fn, t < r1);
    }
  }

  // rows i, i + 1 (i = r0 + 2 tid, even) of block [r0, r1), r1 - r0 even;
  // epi2(i, y_i, y_{i+1}, x_i, x_{i+1}, own)
  template <bool SUB, int K, typename Init, typename Epi2>
  __device__ __forceinline__ static void pairs(const SpmvShape& m, double fn, const double* x,
                                               int r0, int r1, Init& init, Epi2& epi2) {
    gB* xb = (gB*)x;
    gB* sb = (gB*)m.sd;
    const long long S8 = 8 * m.sdg.stride;
    const int t = r0 + 2 * (int)threadIdx.x;
    const int i = min(t, r1 - 2);
    const unsigned ib = 8u * (unsigned)i;          // 16-byte aligned
    const unsigned lastb = 8u * (unsigned)(m.n - 2);
    // every offset first, then all loads: an address register written
    // between the loads can land on a pending load's destination and make
    // the compiler wait for it
    unsigned ou[K], ol[K];
    ou[0] = min(ib + 16, lastb);
    ol[0] = max(ib, 16u) - 16;
#pragma unroll
    for (int k = 1; k < K; ++k) {
      const unsigned ob = 8u * (unsigned)m.sdg.off[k];
      ou[k] = min(ib + ob, lastb);
      ol[k] = max(ib, ob) - ob;
    }
    __builtin_amdgcn_sched_barrier(0);
    dbl2 lo[K], up[K], xl[K], xu[K];
    const dbl2 d = at_nt<dbl2>(sb, ib);
    const dbl2 x0 = at<dbl2>(xb, ib);
    up[0] = at<dbl2>(sb + S8, ib);        // a(i, i+1), a(i+1, i+2)
    lo[0] = at<dbl2>(sb + S8 - 16, ib);   // .y = a(i-1, i)
    xu[0] = at<dbl2>(xb, ou[0]);          // .x = x_{i+2}
    xl[0] = at<dbl2>(xb, ol[0]);          // .y = x_{i-1}
#pragma unroll
    for (int k = 1; k < K; ++k) {
      const unsigned ob = 8u * (unsigned)m.sdg.off[k];
      gB* slot = sb + (k + 1) * S8;
      up[k] = at<dbl2>(slot, ib);
      lo[k] = at<dbl2>(slot - ob, ib);
      xu[k] = at<dbl2>(xb, ou[k]);
      xl[k] = at<dbl2>(xb, ol[k]);
    }
    __builtin_amdgcn_sched_barrier(0);
    double a0 = init(i);
    a0 = term<SUB>(a0, d.x, fn, x0.x);
#pragma unroll
    for (int k = K - 1; k >= 1; --k) a0 = term<SUB>(a0, lo[k].x, fn, xl[k].x);
    a0 = term<SUB>(a0, lo[0].y, fn, xl[0].y);
    a0 = term<SUB>(a0, up[0].x, fn, x0.y);
#pragma unroll
    for (int k = 1; k < K; ++k) a0 = term<SUB>(a0, up[k].x, fn, xu[k].x);
    double a1 = init(i + 1);
    a1 = term<SUB>(a1, d.y, fn, x0.y);
#pragma unroll
    for (int k = K - 1; k >= 1; --k) a1 = term<SUB>(a1, lo[k].y, fn, xl[k].y);
    a1 = term<SUB>(a1, up[0].x, fn, x0.x);
    a1 = term<SUB>(a1, up[0].y, fn, xu[0].x);
#pragma unroll
    for (int k = 1; k < K; ++k) a1 = term<SUB>(a1, up[k].y, fn, xu[k].y);
    epi2(i, a0, a1, x0.x, x0.y, t < r1);
  }

  template <bool SUB, int K, int MODE, typename Init, typename Epi, typename Epi2>
  __device__ __forceinline__ static void run(const SpmvShape& m, double fn, const double* x,
                                             Init init, Epi epi, Epi2 epi2) {
    const XcdWalk w(m.nsorder);
    if (w.first >= w.end) return;
    int2 b = ids(m, w.first);
    for (int lb = w.first; lb < w.end; lb += w.step) {
      const int2 nb = ids(m, min(lb + w.step, w.end - 1));  // in flight during this block
      if constexpr (MODE == 1) pairs<SUB, K>(m, fn, x, b.x, b.y, init, epi2);
      else rows<SUB, K, MODE == 2 ? 2 : 1>(m, fn, x, b.x, b.y, init, epi);
      b = nb;
    }
  }
};

// FP32 row-excess product on the symmetric storage of the FP32 sweeps of
// iterative refinement: slot 0 of sd32 holds s_i = sum_j a_ij, slot k + 1
// the upper diagonal a(i, i + d_k), zero where row i has no such entry; the
// slot stride S leaves at least the largest offset + 4 zeros after every
// slot (lower entries of the first rows read them).  Row i sums
//   s_i x_i,
//   a(i - d_k, i) (x_{i - d_k} - x_i) for k = K-1 .. 0 (the mirror),
//   a(i, i + d_k) (x_{i + d_k} - x_i) for k = 0 .. K-1,
// the CSR order of the row, so the products equal the CSR engines' in the
// row-excess form bit for bit (a missing entry adds a zero term; its x index
// is clamped into the vector).  Modes: 0 one row per thread, blocks of at
// most 256 rows; 1 four rows per thread with 16-byte loads (d_0 = 1, the
// other offsets multiples of 4, n a multiple of 4; blocks of at most 1024
// rows at multiples of 4): rows i .. i + 3 read every term as one aligned
// quadruple, the +-1 neighbours from the quadruples at i - 4 and i + 4.
// epi(i, y_i, x_i, own), epi4(i, y[4], x[4], own).
struct SymEngine32 {
  typedef float flt4 __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) const char gB;
  template <typename V>
  __device__ __forceinline__ static V at(gB* base, unsigned byte) {
    return *(__attribute__((address_space(1))) const V*)(base + byte);
  }
  template <typename V>  // non-temporal: no other row reads it
  __device__ __forceinline__ static V at_nt(gB* base, unsigned byte) {
    return __builtin_nontemporal_load((__attribute__((address_space(1))) const V*)(base + byte));
  }
  __device__ __forceinline__ static float term(float acc, float a, float xj, float xi) {
    return acc + a * (xj - xi);
  }

  template <int K, typename Epi>
  __device__ __forceinline__ static void rows(const SymDiag& g, const float* sd, long long S, int n,
                                              const float* x, int r0, int r1, Epi& epi) {
    gB* xb = (gB*)x;
    gB* sb = (gB*)sd;
    const long long S4 = 4 * S;
    const unsigned lastb = 4u * (unsigned)(n - 1);
    const int t = r0 + (int)threadIdx.x;
    const int i = min(t, r1 - 1);
    const unsigned ib = 4u * (unsigned)i;
    float lo[K], up[K], xl[K], xu[K];
    const float s = at_nt<float>(sb, ib);
    const float x0 = at<float>(xb, ib);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const unsigned ob = 4u * (unsigned)g.off[k];
      gB* slot = sb + (k + 1) * S4;
      up[k] = at<float>(slot, ib);
      lo[k] = at<float>(slot - ob, ib);  // in the zeros after slot k for i < d_k
      xu[k] = at<float>(xb, min(ib + ob, lastb));
      xl[k] = at<float>(xb, max(ib, ob) - ob);
    }
    float acc = s * x0;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) acc = term(acc, lo[k], xl[k], x0);
#pragma unroll
    for (int k = 0; k < K; ++k) acc = term(acc, up[k], xu[k], x0);
    epi(i, acc, x0, t < r1);
  }

  template <int K, typename Epi4>
  __device__ __forceinline__ static void quads(const SymDiag& g, const float* sd, long long S,
                                               int n, const float* x, int r0, int r1, Epi4& epi4) {
    gB* xb = (gB*)x;
    gB* sb = (gB*)sd;
    const long long S4 = 4 * S;
    const int t = r0 + 4 * (int)threadIdx.x;
    const int i = min(t, r1 - 4);
    const unsigned ib = 4u * (unsigned)i;           // 16-byte aligned
    const unsigned lastb = 4u * (unsigned)(n - 4);
    // every offset first, then all loads (as in SymEngine::pairs)
    unsigned ou[K], ol[K];
    ou[0] = min(ib + 16, lastb);
    ol[0] = max(ib, 16u) - 16;
#pragma unroll
    for (int k = 1; k < K; ++k) {
      const unsigned ob = 4u * (unsigned)g.off[k];
      ou[k] = min(ib + ob, lastb);
      ol[k] = max(ib, ob) - ob;
    }
    __builtin_amdgcn_sched_barrier(0);
    flt4 lo[K], up[K], xl[K], xu[K];
    const flt4 s = at_nt<flt4>(sb, ib);
    const flt4 x0 = at<flt4>(xb, ib);
    up[0] = at<flt4>(sb + S4, ib);        // a(i, i+1) .. a(i+3, i+4)
    lo[0] = at<flt4>(sb + S4 - 16, ib);   // .w = a(i-1, i)
    xu[0] = at<flt4>(xb, ou[0]);          // .x = x_{i+4}
    xl[0] = at<flt4>(xb, ol[0]);          // .w = x_{i-1}
#pragma unroll
    for (int k = 1; k < K; ++k) {
      const unsigned ob = 4u * (unsigned)g.off[k];
      gB* slot = sb + (k + 1) * S4;
      up[k] = at<flt4>(slot, ib);
      lo[k] = at<flt4>(slot - ob, ib);
      xu[k] = at<flt4>(xb, ou[k]);
      xl[k] = at<flt4>(xb, ol[k]);
    }
    __builtin_amdgcn_sched_barrier(0);
    float y[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float acc = s[c] * x0[c];
#pragma unroll
      for (int k = K - 1; k >= 1; --k) acc = term(acc, lo[k][c], xl[k][c], x0[c]);
      // offset 1: the row's lower neighbour, then its upper one
      if (c == 0) acc = term(acc, lo[0][3], xl[0][3], x0[0]);
      else acc = term(acc, up[0][c - 1], x0[c - 1], x0[c]);
      if (c == 3) acc = term(acc, up[0][3], xu[0][0], x0[3]);
      else acc = term(acc, up[0][c], x0[c + 1], x0[c]);
#pragma unroll
      for (int k = 1; k < K; ++k) acc = term(acc, up[k][c], xu[k][c], x0[c]);
      y[c] = acc;
    }
    epi4(i, y, x0, t < r1);
  }

  template <int K, int MODE, typename Epi, typename Epi4>
  __device__ __forceinline__ static void run(const SymDiag& g, const float* sd, long long S, int n,
                                             const int* sorder, int nsorder, const float* x,
                                             Epi epi, Epi4 epi4) {
    typedef __attribute__((address_space(4))) const int cint;
    const XcdWalk w(nsorder);
    for (int lb = w.first; lb < w.end; lb += w.step) {
      const int r0 = ((cint*)sorder)[2 * lb], r1 = ((cint*)sorder)[2 * lb + 1];
      if constexpr (MODE == 1) quads<K>(g, sd, S, n, x, r0, r1, epi4);
      else rows<K>(g, sd, S, n, x, r0, r1, epi);
    }
  }
};

// One given row block lb per call (the batched solver's one-block-per-
// workgroup launches): StreamEngine staging without the pipelining.
template <int CAP, typename T>
struct OneBlock {
  using SE = StreamEngine<CAP, false, T>;
  using Smem = typename SE::Smem;
  template <bool SUB, bool LAP = false, typename Init, typename Epi>
  __device__ __forceinline__ static void one(Smem& sm, const SpmvShape& m,
                                             const T* __restrict__ val, T fn, const T* x,
                                             double* s_tmp, double* s_out, int lb, Init init,
                                             Epi epi) {
    constexpr int EPV = SE::EPV;
    const int tid = threadIdx.x;
    const int nnz = m.row[m.n];
    typename SE::Regs g;
    SE::fetch(m, val, nnz, lb, g);
    const int r0 = g.r0, nr = g.nr, s = g.s, e = g.e, rs = g.rs, re = g.re;
    if (e - s > CAP) {
      long_row<SUB, LAP>(m, val, fn, x, r0, s, e, s_tmp, s_out, init, epi);
      return;
    }
    g.v.put(sm.val, nnz, s, e);
    g.c.put(sm.col, nnz, s, e);
    __syncthreads();
    const int ov = s & ~(EPV - 1), oc = s & ~(SE::EPC - 1);
    if (tid < nr) {
      const int i = r0 + tid;
      const T xi = x[sm.col[rs - oc]];
      T acc = init(i);
      for (int j = rs; j < re; ++j) {
        const T t = row_term<LAP>(sm.val[j - ov] * fn, x[sm.col[j - oc]], xi, j == rs);
        acc = SUB ? acc - t : acc + t;
      }
      epi(i, acc, xi, sm.val[rs - ov] * fn);
    }
    __syncthreads();
  }
};

// Stage engine with element-wise staging loads, for a val array that is not
// 16-byte aligned or a matrix with fewer than 4 nonzeros.
template <int CAP, typename T>
struct ScalarStage {
  struct Smem {
    T val[CAP];
    int col[CAP];
  };

  template <bool SUB, bool LAP = false, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem& sm, const SpmvShape& m,
                                             const T* __restrict__ val, T fn,
                                             const T* __restrict__ x, double* s_tmp,
                                             double* s_out, Init init, Epi epi) {
    const int tid = threadIdx.x;
    const XcdWalk w(m.nblk);
    for (int lb = w.first; lb < w.end; lb += w.step) {
      const int r0 = m.blk[lb], nr = m.blk[lb + 1] - r0;
      const int s = m.row[r0], e = m.row[r0 + nr];
      if (e - s > CAP) {
        long_row<SUB, LAP>(m, val, fn, x, r0, s, e, s_tmp, s_out, init, epi);
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
        const T xi = x[sm.col[rs - s]];
        T acc = init(i);
        for (int j = rs; j < re; ++j) {
          const T t = row_term<LAP>(sm.val[j - s] * fn, x[sm.col[j - s]], xi, j == rs);
          acc = SUB ? acc - t : acc + t;
        }
        epi(i, acc, xi, sm.val[rs - s] * fn);
      }
      __syncthreads();
    }
  }

  // one given row block lb (see OneBlock)
  template <bool SUB, bool LAP = false, typename Init, typename Epi>
  __device__ __forceinline__ static void one(Smem& sm, const SpmvShape& m,
                                             const T* __restrict__ val, T fn,
                                             const T* __restrict__ x, double* s_tmp,
                                             double* s_out, int lb, Init init, Epi epi) {
    const int tid = threadIdx.x;
    const int r0 = m.blk[lb], nr = m.blk[lb + 1] - r0;
    const int s = m.row[r0], e = m.row[r0 + nr];
    if (e - s > CAP) {
      long_row<SUB, LAP>(m, val, fn, x, r0, s, e, s_tmp, s_out, init, epi);
      return;
    }
    for (int t = tid; t < e - s; t += kBlock) {
      sm.val[t] = val[s + t];
      sm.col[t] = m.col[s + t];
    }
    __syncthreads();
    if (tid < nr) {
      const int i = r0 + tid;
      const int rs = m.row[i], re = m.row[i + 1];
      const T xi = x[sm.col[rs - s]];
      T acc = init(i);
      for (int j = rs; j < re; ++j) {
        const T t = row_term<LAP>(sm.val[j - s] * fn, x[sm.col[j - s]], xi, j == rs);
        acc = SUB ? acc - t : acc + t;
      }
      epi(i, acc, xi, sm.val[rs - s] * fn);
    }
    __syncthreads();
  }
};

// ---------------------------------------------------------------------------
__device__ __forceinline__ double dpp_shr1(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xf, 0xf, false);  // row_shr:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float dpp_shr1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
}

template <int L, typename T>
struct WaveEngine {
  static_assert(L == 4 || L == 8 || L == 16, "lanes per row must divide a DPP row");
  struct Smem {
    int unused;
  };

  template <bool SUB, typename Init, typename Epi>
  __device__ __forceinline__ static void run(Smem&, const SpmvShape& m,
                                             const T* __restrict__ val, T fn,
                                             const T* __restrict__ x, double*, double*,
                                             Init init, Epi epi) {
    const int jl = threadIdx.x & (L - 1);
    const int gpb = kBlock / L;  // row groups per workgroup
    const XcdWalk w(m.n, gpb);
    for (int i0 = w.first; i0 < w.end; i0 += w.step) {
      const int i = i0 + (int)threadIdx.x / L;
      if (i >= w.end) continue;  // whole groups drop out together
      const int rs = m.row[i], re = m.row[i + 1];
      T acc = jl == 0 ? init(i) : T(0), xi = T(0), di = T(0);
      for (int j0 = rs; j0 < re; j0 += L) {
        const int j = j0 + jl;
        const int jc = min(j, re - 1);  // loads never branch; terms past the end are 0
        const T a = val[jc] * fn, xv = x[m.col[jc]];
        const T t = j < re ? a * xv : T(0);
        if (j == rs) {
          xi = xv;
          di = a;
        }
        T sacc = SUB ? acc - t : acc + t;  // meaningful on lane 0
#pragma unroll
        for (int k = 1; k < L; ++k) {
          const T prev = dpp_shr1(sacc);
          if (jl == k) sacc = SUB ? prev - t : prev + t;
        }
        // carry = the group's last lane (terms past the row end added +-0)
        acc = __shfl(sacc, ((int)threadIdx.x & 63) | (L - 1), 64);
      }
      if (jl == 0) epi(i, acc, xi, di);
    }
  }
};

// kind: 0 stage/kCap, 1 stage/kCap2, 2 wave/4, 3 wave/8, 4 wave/16,
// 5 pipelined stage/kCap, 6 the same with offset-coded columns, 7 the same
// with row patterns (also kind 8: the symmetric engine serves only the CG
// product and the plain product, when Launch::sym is set; every other
// kernel uses the row-pattern engine).  The stage engines fall back to scalar staging
// (ScalarStage) when val is not 16-byte aligned or nnz < 4.
#define TFG_ENGINE_SWITCH_T(L, T, CALL)                                            \
  do {                                                                             \
    switch ((L).kind) {                                                            \
      case 0:                                                                      \
        if ((L).vec_val) { using E = StreamEngine<kCap, false, T>; CALL; }         \
        else { using E = ScalarStage<kCap, T>; CALL; }                             \
        break;                                                                     \
      case 1:                                                                      \
        if ((L).vec_val) { using E = StreamEngine<kCap2, false, T>; CALL; }         \
        else { using E = ScalarStage<kCap2, T>; CALL; }                             \
        break;                                                                     \
      case 5:                                                                      \
        if ((L).vec_val) { using E = StreamEngine<kCap, true, T>; CALL; }          \
        else { using E = ScalarStage<kCap, T>; CALL; }                             \
        break;                                                                     \
      case 6:                                                                      \
        if ((L).vec_val) { using E = StreamEngine<kCap, true, T, true>; CALL; }    \
        else { using E = ScalarStage<kCap, T>; CALL; }                             \
        break;                                                                     \
      case 7:                                                                      \
      case 8:                                                                      \
        if ((L).vec_val) { using E = PatternEngine<kCap, true, T>; CALL; }         \
        else { using E = ScalarStage<kCap, T>; CALL; }                             \
        break;                                                                     \
      case 2: { using E = WaveEngine<4, T>; CALL; } break;                         \
      case 3: { using E = WaveEngine<8, T>; CALL; } break;                         \
      default: { using E = WaveEngine<16, T>; CALL; } break;                       \
    }                                                                              \
  } while (0)

}  // namespace tfg
