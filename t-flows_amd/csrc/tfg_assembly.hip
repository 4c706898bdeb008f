// Pressure-system assembly on the device from face data, so that a pressure
// solve needs no host-to-device copy of A % val and b.  What it computes is
// the reference's, in the reference's order:
//   A % val = 0, b = 0                          Compute_Pressure.f90:121-122
//   internal faces s (in order): a12 = fc(s) (f(s) v_m(c1) + (1 - f(s)) v_m(c2)),
//     val(pos(1,s)) = val(pos(2,s)) = -a12, val(dia(c1)) += a12,
//     val(dia(c2)) += a12                          Rhie_And_Chow.f90:184-205
//   all faces s (in order): b(c1) -= flux(s), b(c2) += flux(s) (internal);
//     boundary faces with a pressure / ambient condition:
//     val(dia(c1)) += fc(s) v_m(c1)             Compute_Pressure.f90:147-169
//   b(c1) = 0 in pressure regions                Compute_Pressure.f90:173-178
//   optionally a_ii *= (1 + MICRO)              Solver_Mod/Set_Singular.f90:28-31
// The face loop scatters into the diagonal and b, so it is turned around:
// one thread per row walks, in increasing face order, the faces of its cell
// (a list built once per structure) and forms its diagonal and b entry as
// the serial loop would; each off-diagonal entry takes -a12 of the last face
// that sets it.  Every value is bit-identical to the serial assembly.
#include "tfg_spmv.h"

namespace tfg {

// a12 of internal face s                             Rhie_And_Chow.f90:200
__device__ __forceinline__ double face_a12(const AsmArgs& g, int s) {
  const double fs = g.fw[s];
  return g.fc[s] * (fs * g.vm[g.c1[s]] + (1.0 - fs) * g.vm[g.c2[s]]);
}

__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs g) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= g.n) return;
  const int fs0 = g.cf_off[c], fs1 = g.cf_off[c + 1];
  // diagonal: internal faces in face order, then the boundary faces'
  // terms in face order (the second loop of Compute_Pressure)
  double diag = 0.0;
  for (int t = fs0; t < fs1; ++t) {
    const int s = g.cf_face[t] >> 1;
    if (g.c2[s] >= 0) diag = diag + face_a12(g, s);
  }
  double bc = 0.0;
  bool zero = false;
  for (int t = fs0; t < fs1; ++t) {
    const int e = g.cf_face[t], s = e >> 1;
    const double flux = g.vflux[s];
    if (g.c2[s] >= 0) {
      bc = (e & 1) ? bc + flux : bc - flux;
    } else {
      bc = bc - flux;
      if (g.flags[s] & 1) diag = diag + g.fc[s] * g.vm[c];
      if (g.flags[s] & 2) zero = true;
    }
  }
  if (g.singular) diag = diag * (1.0 + 1.0e-6);
  const int rs = g.row[c], re = g.row[c + 1];
  g.val[rs] = diag;
  for (int j = rs + 1; j < re; ++j) {
    const int s = g.ent_face[j];
    g.val[j] = s < 0 ? 0.0 : -face_a12(g, s);
  }
  g.b[c] = zero ? 0.0 : bc;
}

hipError_t launch_assemble(hipStream_t stream, const AsmArgs& g) {
  k_assemble<<<(g.n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(g);
  return hipGetLastError();
}

}  // namespace tfg
