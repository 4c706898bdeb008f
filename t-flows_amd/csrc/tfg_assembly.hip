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
//   boundary regions: b(c1) = 0 in pressure regions, b(c1) = fc(s) v_m(c1)
//     pp(c2) in ambient regions, assigned region by region and face by
//     face                                        Compute_Pressure.f90:173-188
//   optionally a_ii *= (1 + MICRO)              Solver_Mod/Set_Singular.f90:28-31
// The face loop scatters into the diagonal and b, so it is turned around:
// one thread per row walks, in increasing face order, the faces of its cell
// (a list built once per structure) and forms its diagonal and b entry as
// the serial loop would; each off-diagonal entry takes -a12 of the last face
// that sets it.  The region assignments of b follow face order, which is the
// region loop's order when each region's faces are one range and the ranges
// ascend with the region number (Faces_In_Region, as T-Flows numbers its
// boundary faces).  Every value is bit-identical to the serial assembly.
#include "tflows_gpu.h"
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
  double bc = 0.0, bset = 0.0;
  bool set = false;
  for (int t = fs0; t < fs1; ++t) {
    const int e = g.cf_face[t], s = e >> 1;
    const double flux = g.vflux[s];
    if (g.c2[s] >= 0) {
      bc = (e & 1) ? bc + flux : bc - flux;
    } else {
      bc = bc - flux;
      const int fl = g.flags[s];
      if (fl & 1) diag = diag + g.fc[s] * g.vm[c];
      if (fl & 2) {
        bset = 0.0;
        set = true;
      } else if (fl & 4) {
        const double a12 = g.fc[s] * g.vm[c];
        bset = a12 * g.pp[g.c2b[s] - 1];
        set = true;
      }
    }
  }
  if (g.singular) diag = diag * (1.0 + 1.0e-6);
  const int rs = g.row[c], re = g.row[c + 1];
  g.val[rs] = diag;
  for (int j = rs + 1; j < re; ++j) {
    const int s = g.ent_face[j];
    g.val[j] = s < 0 ? 0.0 : -face_a12(g, s);
  }
  g.b[c] = set ? bset : bc;
}

hipError_t launch_assemble(hipStream_t stream, const AsmArgs& g) {
  k_assemble<<<(g.n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(g);
  return hipGetLastError();
}

// vel_max = max(MICRO, max_c sqrt(u_c^2 + v_c^2 + w_c^2))  Compute_Momentum.f90:201-204
// (a maximum: the same bits in any order).  One workgroup per chunk of
// cells, the chunk maxima combined by the second kernel.
__global__ __launch_bounds__(kBlock) void k_vel_max(int n, const double* __restrict__ u,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ part) {
  __shared__ double s_tmp[4], s_out[1];
  double m[1] = {1.0e-6};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    m[0] = fmax(m[0], sqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]));
  const int ops[1] = {OP_MAX};
  publish<1>(m, ops, part, 1, s_tmp, s_out);
}

__global__ __launch_bounds__(kBlock) void k_vel_max_final(const double* __restrict__ part,
                                                          int G, double* out) {
  __shared__ double s_tmp[4], s_out[1];
  const int ops[1] = {OP_MAX};
  gather<1>(part, 1, G, ops, s_tmp, s_out);
  if (threadIdx.x == 0) out[0] = s_out[0];
}

// p_min, p_max over the cells (Compute_Pressure.f90:104-109; HUGE
// starts, which a finite field always replaces)
__global__ __launch_bounds__(kBlock) void k_range(int n, const double* __restrict__ v,
                                                 double* __restrict__ part) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  double m[2] = {kHuge, -kHuge};
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    m[0] = fmin(m[0], v[i]);
    m[1] = fmax(m[1], v[i]);
  }
  const int ops[2] = {OP_MIN, OP_MAX};
  publish<2>(m, ops, part, 1024, s_tmp, s_out);
}

__global__ __launch_bounds__(kBlock) void k_range_final(const double* __restrict__ part, int G,
                                                       double* out) {
  __shared__ double s_tmp[2 * 4], s_out[2];
  const int ops[2] = {OP_MIN, OP_MAX};
  gather<2>(part, 1024, G, ops, s_tmp, s_out);
  if (threadIdx.x < 2) out[threadIdx.x] = s_out[threadIdx.x];
}

hipError_t launch_range(hipStream_t stream, int n, const double* v, double* part, int G,
                        double* out) {
  if (G > 1024) return hipErrorInvalidValue;
  k_range<<<G, kBlock, 0, stream>>>(n, v, part);
  k_range_final<<<1, kBlock, 0, stream>>>(part, G, out);
  return hipGetLastError();
}

hipError_t launch_vel_max(hipStream_t stream, int n, const double* u, const double* v,
                          const double* w, double* part, int G, double* out) {
  k_vel_max<<<G, kBlock, 0, stream>>>(n, u, v, w, part);
  k_vel_max_final<<<1, kBlock, 0, stream>>>(part, G, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Momentum assembly of one velocity component (Compute_Momentum.f90:119-366
// with Numerics_Mod).  As for the pressure, every face loop of the reference
// is turned around: one thread per cell walks its faces and sums its own
// entries in the reference's order.  That order is increasing face index in
// every loop except the advection loops, which take the boundary faces first
// and then the domain faces (Advection_Term.f90:35-102: boundary regions,
// then Faces_In_Domain_And_At_Buffers).  The two coincide for T-Flows' face
// numbering (boundary regions first) and are followed literally otherwise.
// A face's value is computed by both of its cells with the same arithmetic,
// so both see the same bits.
// ---------------------------------------------------------------------------

namespace {

__device__ __forceinline__ double cell(const double* a, int c) { return a[c - 1]; }
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return b > a ? b : a; }

// min / max of phi over a cell and its face neighbours (Min_Max.f90:19-34);
// a boundary cell has one face, to cell c1 of that face
__device__ __forceinline__ void minmax_at(const MomArgs& m, int C, int s, double& mn, double& mx) {
  if (C > 0) {
    mn = m.pmin[C - 1];
    mx = m.pmax[C - 1];
  } else {
    const double a = cell(m.phi, C), o = cell(m.phi, m.c1[s] + 1);
    mn = dmin(a, o);
    mx = dmax(a, o);
  }
}

// face value of face s for every scheme but CENTRAL
// (Numerics_Mod/Advection_Scheme.f90:22-183)
__device__ double adv_scheme(const MomArgs& m, int s, int c1, int c2) {
  const double* phi = m.phi;
  const double F = m.v_flux[s];
  double fj, sgn;
  int c, d;
  if (F > 0.0) {
    fj = 1.0 - m.f[s]; c = c1; d = c2; sgn = +1.0;
  } else {
    fj = m.f[s]; c = c2; d = c1; sgn = -1.0;
  }
  if (m.adv == TFG_ADV_UPWIND) return cell(phi, c);
  if (m.adv == TFG_ADV_BLENDED) {
    const double phi_f_c = cell(phi, c) + fj * sgn * (cell(phi, c2) - cell(phi, c1));
    const double phi_f_u = cell(phi, c);
    return m.blend * phi_f_c + (1.0 - m.blend) * phi_f_u;
  }
  const double gd = cell(m.g[0], c) * m.dv[0][s] + cell(m.g[1], c) * m.dv[1][s] +
                    cell(m.g[2], c) * m.dv[2][s];
  if (m.adv == TFG_ADV_GAMMA) {
    const double beta = 0.1;
    const double phi_star = F > 0.0 ? 1.0 - (cell(phi, d) - cell(phi, c)) / (2.0 * gd)
                                    : 1.0 + (cell(phi, d) - cell(phi, c)) / (2.0 * gd);
    const double gamma_c = phi_star / beta;
    if (phi_star < beta && phi_star > 0.0)
      return (1.0 - gamma_c * (1.0 - m.f[s])) * cell(phi, c) + gamma_c * (1.0 - m.f[s]) * cell(phi, d);
    if (phi_star < 1.0 && phi_star >= beta) return m.f[s] * cell(phi, c) + (1.0 - m.f[s]) * cell(phi, d);
    return cell(phi, c);
  }
  const double phi_star = F > 0.0 ? cell(phi, d) - 2.0 * gd : cell(phi, d) + 2.0 * gd;
  double mn, mx;
  minmax_at(m, c, s, mn, mx);
  const double phi_u = dmax(mn, dmin(phi_star, mx));
  double denom = cell(phi, d) - cell(phi, c);
  if (fabs(denom) < 1.0e-15) denom = denom + copysign(1.0e-15, denom);
  const double rj = (cell(phi, c) - phi_u) / denom;
  const double g_d = 0.5 * fj * (1.0 + fj), g_u = 0.5 * fj * (1.0 - fj);
  double phij = fj;
  if (m.adv == TFG_ADV_QUICK) {
    const double alfa = 0.0;
    phij = (g_d - alfa) + (g_u + alfa) * rj;
  } else if (m.adv == TFG_ADV_LUDS) {
    const double alfa = 0.5 * fj * (1 + fj);
    phij = (g_d - alfa) + (g_u + alfa) * rj;
  } else if (m.adv == TFG_ADV_MINMOD) {
    phij = fj * dmax(0.0, dmin(rj, 1.0));
  } else if (m.adv == TFG_ADV_SMART) {
    const double beta1 = 3.0, beta2 = 1.0;
    phij = dmax(0.0, dmin(dmin((beta1 - 1.0) * rj, g_d + g_u * rj), beta2));
  } else if (m.adv == TFG_ADV_AVL_SMART) {
    const double beta1 = 1.0 + fj * (2.0 + fj), beta2 = fj * (2.0 - fj);
    phij = dmax(0.0, dmin(dmin((beta1 - 1.0) * rj, g_d + g_u * rj), beta2));
  } else if (m.adv == TFG_ADV_SUPERBEE) {
    phij = 0.5 * dmax(dmax(0.0, dmin(2.0 * rj, 1.0)), dmin(rj, 2.0));
  }
  return cell(phi, c) + phij * sgn * (cell(phi, c2) - cell(phi, c1));
}

// m12 and m21 of face s (Compute_Momentum.f90:271-282)
__device__ __forceinline__ void face_m(const MomArgs& m, int s, int C1, int C2, double& m12,
                                       double& m21) {
  const double m0 = m.vis_eff[s] * m.fc[s];
  m12 = m0;
  m21 = m0;
  if (m.blend_matrix) {
    m12 = m12 - dmin(m.v_flux[s], 0.0) * cell(m.rho, C1);
    m21 = m21 + dmax(m.v_flux[s], 0.0) * cell(m.rho, C2);
  }
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_mom_minmax(MomArgs m) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.n) return;
  const int C = c + 1;
  double mn = cell(m.phi, C), mx = mn;
  for (int t = m.cf_off[c]; t < m.cf_off[c + 1]; ++t) {
    const int e = m.cf_face[t], s = e >> 1;
    const double o = (e & 1) ? cell(m.phi, m.c1[s] + 1) : cell(m.phi, m.c2b[s]);
    mn = dmin(mn, o);
    mx = dmax(mx, o);
  }
  m.pmin[c] = mn;
  m.pmax[c] = mx;
}

__global__ __launch_bounds__(kBlock) void k_momentum(MomArgs m) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.n) return;
  const int C = c + 1;
  const int t0 = m.cf_off[c], t1 = m.cf_off[c + 1];
  const double rho_c = cell(m.rho, C);
  // advection, b = fi (Advection_Term.f90:35-120): boundary faces, then
  // domain faces, each in increasing order
  double adv = 0.0, upw = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int t = t0; t < t1; ++t) {
      const int e = m.cf_face[t], s = e >> 1;
      const int C1 = m.c1[s] + 1, C2 = m.c2b[s];
      if ((pass == 0) != (C2 < 0)) continue;
      const double F = m.v_flux[s];
      double phif = m.f[s] * cell(m.phi, C1) + (1.0 - m.f[s]) * cell(m.phi, C2);
      if (m.adv != TFG_ADV_CENTRAL) phif = adv_scheme(m, s, C1, C2);
      adv = (e & 1) ? adv + F * phif * rho_c : adv - F * phif * rho_c;
      if (m.blend_matrix) {
        const double up = F < 0 ? cell(m.phi, C2) : cell(m.phi, C1);
        upw = (e & 1) ? upw + F * up * rho_c : upw - F * up * rho_c;
      }
    }
  }
  double fi = 0.0;
  fi = fi + adv - upw;
  // diffusion (Compute_Momentum.f90:238-301)
  const int i = m.comp;
  const double *ui_i = m.g[(i - 1) % 3], *ui_j = m.g[i % 3], *ui_k = m.g[(i + 1) % 3];
  const double *si = m.sv[(i - 1) % 3], *sj = m.sv[i % 3], *sk = m.sv[(i + 1) % 3];
  const double *di = m.dv[(i - 1) % 3], *dj = m.dv[i % 3], *dk = m.dv[(i + 1) % 3];
  double cross = 0.0, diag = 0.0;
  for (int t = t0; t < t1; ++t) {
    const int e = m.cf_face[t], s = e >> 1;
    const int C1 = m.c1[s] + 1, C2 = m.c2b[s];
    const double w = m.fw[s], vis_eff = m.vis_eff[s];
    const double f_stress = m.f_stress ? m.f_stress[s] : 0.0;
    const double ui_i_f = w * cell(ui_i, C1) + (1.0 - w) * cell(ui_i, C2);
    const double ui_j_f = w * cell(ui_j, C1) + (1.0 - w) * cell(ui_j, C2);
    const double ui_k_f = w * cell(ui_k, C1) + (1.0 - w) * cell(ui_k, C2);
    const double uj_i_f = w * cell(m.uj_i, C1) + (1.0 - w) * cell(m.uj_i, C2);
    const double uk_i_f = w * cell(m.uk_i, C1) + (1.0 - w) * cell(m.uk_i, C2);
    const double ui_si = (ui_i_f + ui_i_f) * si[s] + (ui_j_f + uj_i_f) * sj[s] + (ui_k_f + uk_i_f) * sk[s];
    const double ui_di = ui_i_f * di[s] + ui_j_f * dj[s] + ui_k_f * dk[s];
    const double f_ex = vis_eff * ui_si;
    const double m0 = vis_eff * m.fc[s];
    const double f_im = ui_di * m0;
    if (e & 1) cross = cross - f_ex + f_im - f_stress * rho_c;
    else cross = cross + f_ex - f_im + f_stress * rho_c;
    double m12, m21;
    face_m(m, s, C1, C2, m12, m21);
    if (C2 > 0) {
      diag = diag + ((e & 1) ? m21 : m12);
    } else if (m.wall[s]) {
      diag = diag + m12;
      fi = fi + m12 * cell(m.phi, C2);
    }
  }
  fi = fi + cross;                                        // :306-308
  const double vol = cell(m.vol, C);
  const double a0 = rho_c * vol / m.dt;                   // Inertial_Term.f90:27-41
  if (m.td == TFG_TD_LINEAR) {
    diag = diag + a0;
    fi = fi + a0 * cell(m.phi_o, C);
  } else if (m.td == TFG_TD_PARABOLIC) {
    diag = diag + 1.5 * a0;
    fi = fi + 2.0 * a0 * cell(m.phi_o, C) - 0.5 * a0 * cell(m.phi_oo, C);
  }
  fi = fi + m.p_drop * vol;                               // :322-326
  const double cf = m.cell_f ? cell(m.cell_f, C) : 0.0;
  fi = fi + cf * vol;                                     // :328-331
  double bb = fi - cell(m.p_i, C) * vol;                  // :337-340
  if (m.st) bb = bb + cell(m.st, C) * vol;                // :345-350
  m.v_m[c] = vol / diag;                                  // :361-363
  diag = diag / m.urf;                                    // Under_Relax.f90:18-24
  bb = bb + diag * (1.0 - m.urf) * cell(m.phi, C);
  // row c: the diagonal first, then each off-diagonal entry from its face
  const int rs = m.row[c], re = m.row[c + 1];
  m.val[rs] = diag;
  for (int j = rs + 1; j < re; ++j) {
    const int s = m.ent_face[j];
    if (s < 0) {
      m.val[j] = 0.0;
    } else {
      const int C1 = m.c1[s] + 1, C2 = m.c2b[s];
      double m12, m21;
      face_m(m, s, C1, C2, m12, m21);
      m.val[j] = 0.0 - (C1 == C ? m12 : m21);
    }
  }
  m.b[c] = bb;
}

// ---------------------------------------------------------------------------
// Rhie-Chow volume fluxes (Rhie_And_Chow.f90:86-259), one thread per face.
// The cell quantities (the corrected velocities u_c, the pressure terms pst)
// are formed for each of the face's two cells with the serial loop's
// operations, so both faces of a cell see the same bits.
// ---------------------------------------------------------------------------
namespace {

struct RcCell {
  double u, v, w, px, py, pz, tm;
};

__device__ __forceinline__ RcCell rc_cell(const RcArgs& a, int c) {
  const tfg_rhie_chow& r = a.r;
  RcCell o;
  o.px = r.px[c];
  o.py = r.py[c];
  o.pz = r.pz[c];
  if (a.vof) {                                            // :107-112
    o.px = o.px - r.sigma * r.curv[c] * r.col_x[c];
    o.py = o.py - r.sigma * r.curv[c] * r.col_y[c];
    o.pz = o.pz - r.sigma * r.curv[c] * r.col_z[c];
  }
  o.u = r.u[c];                                           // :130-134
  o.v = r.v[c];
  o.w = r.w[c];
  o.tm = 0.0;
  if (r.choi) {                                           // :139-163
    o.tm = r.v_m[c] * r.density[c] / r.dt;
    o.u = o.u - (r.w_o * r.u_o[c] + r.w_oo * r.u_oo[c]) * o.tm;
    o.v = o.v - (r.w_o * r.v_o[c] + r.w_oo * r.v_oo[c]) * o.tm;
    o.w = o.w - (r.w_o * r.w_o_[c] + r.w_oo * r.w_oo_[c]) * o.tm;
  }
  if (r.gu) {                                             // :169-179
    o.u = o.u - r.v_m[c] * r.cell_fx[c];
    o.v = o.v - r.v_m[c] * r.cell_fy[c];
    o.w = o.w - r.v_m[c] * r.cell_fz[c];
  }
  return o;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_rhie_chow(RcArgs a) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= a.n_faces) return;
  const int c1 = a.c1[s], c2 = a.c2[s];
  if (c2 < 0) return;                                     // domain faces only
  const tfg_rhie_chow& r = a.r;
  const RcCell A = rc_cell(a, c1), B = rc_cell(a, c2);
  const double fs = r.f[s];
  double pst_d = r.p[c1] - r.p[c2];                       // :94-98
  if (a.vof) {                                            // :115-123
    const double curv_f = fs * r.curv[c1] + (1.0 - fs) * r.curv[c2];
    pst_d = pst_d - r.sigma * curv_f * (r.col[c1] - r.col[c2]);
  }
  const double u_f = fs * A.u + (1.0 - fs) * B.u;
  const double v_f = fs * A.v + (1.0 - fs) * B.v;
  const double w_f = fs * A.w + (1.0 - fs) * B.w;
  const double vm1 = r.v_m[c1], vm2 = r.v_m[c2];
  const double a12 = r.fc[s] * (fs * vm1 + (1.0 - fs) * vm2);
  const double px_f = (fs * (A.px * vm1) + (1.0 - fs) * (B.px * vm2)) * r.dx[s];
  const double py_f = (fs * (A.py * vm1) + (1.0 - fs) * (B.py * vm2)) * r.dy[s];
  const double pz_f = (fs * (A.pz * vm1) + (1.0 - fs) * (B.pz * vm2)) * r.dz[s];
  double fl = (u_f * r.sx[s] + v_f * r.sy[s] + w_f * r.sz[s]) + a12 * pst_d +
              r.fc[s] * (px_f + py_f + pz_f);             // :226-230
  if (r.choi)                                             // :235-246
    fl = fl + (fs * A.tm + (1.0 - fs) * B.tm) * (r.w_o * r.flux_o[s] + r.w_oo * r.flux_oo[s]);
  if (r.gu)                                               // :250-256
    fl = fl + (fs * vm1 + (1.0 - fs) * vm2) *
                  (r.face_fx[s] * r.sx[s] + r.face_fy[s] * r.sy[s] + r.face_fz[s] * r.sz[s]);
  a.v_flux[s] = fl;
}

hipError_t launch_rhie_chow(hipStream_t stream, const RcArgs& a) {
  if (a.n_faces <= 0) return hipSuccess;
  k_rhie_chow<<<(a.n_faces + kBlock - 1) / kBlock, kBlock, 0, stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_momentum(hipStream_t stream, const MomArgs& m) {
  const int grid = (m.n + kBlock - 1) / kBlock;
  if (m.adv != TFG_ADV_CENTRAL) {
    k_mom_minmax<<<grid, kBlock, 0, stream>>>(m);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  k_momentum<<<grid, kBlock, 0, stream>>>(m);
  return hipGetLastError();
}

}  // namespace tfg
