// Deterministic reductions for gfx950 (wave64) without atomics or fences.
//
// A producer kernel reduces inside each workgroup (fixed butterfly inside the
// wavefront, fixed-order combination of the four wavefronts) and stores one
// partial per workgroup.  The consumer -- the next kernel in stream order,
// which sees those stores because a kernel boundary separates them -- sweeps
// the partials in a fixed order in EVERY workgroup that needs the total.  All
// workgroups therefore compute bit-identical scalars, and the result depends
// only on the grid shape, never on timing.  (A last-workgroup ticket would
// need a device-scope release per workgroup, which on this multi-XCD part is
// an L2 write-back.)
#pragma once

#include "tfg_internal.h"

namespace tfg {


// Reduction operators passed by value to a kernel.
template <int NV>
struct Ops {
  int v[NV];
};

// Thread 0 of workgroup 0: the one thread that writes a kernel's scalars.
__device__ __forceinline__ bool lead_thread() { return blockIdx.x == 0 && threadIdx.x == 0; }

__device__ __forceinline__ double red_apply(int op, double a, double b) {
  return op == OP_SUM ? a + b : (op == OP_MIN ? fmin(a, b) : fmax(a, b));
}

__device__ __forceinline__ double red_identity(int op) {
  return op == OP_SUM ? 0.0 : (op == OP_MIN ? INFINITY : -INFINITY);
}

// Value of lane (lane ^ OFF) inside the wavefront.  Distances 32 and 16
// cross rows of 16 lanes and go through the LDS crossbar (ds_bpermute);
// 8, 4, 2, 1 stay inside a row and are DPP moves on the VALU: row rotate by
// 8; half-row mirror then quad reverse (lane ^ 7 ^ 3); quad permutes.  Every
// lane of the wavefront must be active.
template <int OFF>
__device__ __forceinline__ int xor_lane_b32(int v) {
  if constexpr (OFF == 8) {
    return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);   // row_ror:8
  } else if constexpr (OFF == 4) {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, false);         // quad_perm 3,2,1,0
  } else if constexpr (OFF == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
  } else {
    static_assert(OFF == 1, "DPP distances are 8, 4, 2, 1");
    return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
  }
}
template <int OFF>
__device__ __forceinline__ double xor_lane(double v) {
  if constexpr (OFF >= 16) {
    return __shfl_xor(v, OFF, 64);
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane_b32<OFF>((int)(unsigned)(b & 0xffffffffLL));
    const int hi = xor_lane_b32<OFF>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}

// One butterfly step on NV values: all NV exchanges, then all NV combines.
template <int OFF, int NV>
__device__ __forceinline__ void butterfly_step(double (&w)[NV], const int (&ops)[NV]) {
  double o[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) o[k] = xor_lane<OFF>(w[k]);
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k] = red_apply(ops[k], w[k], o[k]);
}

// Reduce NV values over the workgroup; on return s_out[k] holds the totals
// (visible to all threads).  s_tmp needs NV * (kBlock/64) doubles.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&ops)[NV],
                                             double* s_tmp, double* s_out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // xor butterfly inside the wavefront (lane distance 32, 16, .., 1), all NV
  // values per step, so a step's NV lane exchanges are in flight together
  double w[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k] = v[k];
  butterfly_step<32>(w, ops);
  butterfly_step<16>(w, ops);
  butterfly_step<8>(w, ops);
  butterfly_step<4>(w, ops);
  butterfly_step<2>(w, ops);
  butterfly_step<1>(w, ops);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s_tmp[k * (kBlock / 64) + wid] = w[k];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double t = s_tmp[k * (kBlock / 64)];
    for (int w = 1; w < kBlock / 64; ++w) t = red_apply(ops[k], t, s_tmp[k * (kBlock / 64) + w]);
    s_out[k] = t;
  }
  __syncthreads();
}

// Producer side: reduce over the workgroup and store the totals at
// part[k*pstride + blockIdx.x].
template <int NV>
__device__ __forceinline__ void publish(double (&v)[NV], const int (&ops)[NV], double* part,
                                        int pstride, double* s_tmp, double* s_out) {
  block_reduce<NV>(v, ops, s_tmp, s_out);
  if (threadIdx.x < NV) part[threadIdx.x * pstride + blockIdx.x] = s_out[threadIdx.x];
}

// Consumer side: combine the G partials of a previous kernel in a fixed order.
// Result in s_out[k] for all threads.
template <int NV>
__device__ __forceinline__ void gather(const double* __restrict__ part, int pstride, int G,
                                       const int (&ops)[NV], double* s_tmp, double* s_out) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = red_identity(ops[k]);
  for (int g = threadIdx.x; g < G; g += kBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = red_apply(ops[k], acc[k], part[k * pstride + g]);
  }
  block_reduce<NV>(acc, ops, s_tmp, s_out);
}

// Compensated (Neumaier) addition: the running sum and the rounding errors
// of its additions, added at the end.
struct CSum {
  double s, c;
  __device__ __forceinline__ void add(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c = c + ((s - t) + v);
    else c = c + ((v - t) + s);
    s = t;
  }
};

// Double-double sums for the one reduction that must not lose digits to
// cancellation (sum_i b_i of a right-hand side that sums to nearly zero):
// hi + lo with |lo| <= ulp(hi) / 2; additions by Knuth's two-sum.
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b.hi - bb);
  const double lo = err + (a.lo + b.lo);
  const double hi = s + lo;
  return DD{hi, lo - (hi - s)};
}
// Sum over the workgroup; on return s_dd[0] holds it for every thread
// (s_dd: kBlock / 64 + 1 entries).
__device__ __forceinline__ DD dd_block(DD v, DD* s_dd) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const DD o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
    v = dd_add(v, o);
  }
  if (lane == 0) s_dd[1 + wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    DD t = s_dd[1];
    for (int w = 1; w < kBlock / 64; ++w) t = dd_add(t, s_dd[1 + w]);
    s_dd[0] = t;
  }
  __syncthreads();
  return s_dd[0];
}

}  // namespace tfg
