/*
 * Synthetic linear systems in T-Flows' matrix layout (host code, plain C).
 *
 * Every generator writes a 1-based CSR matrix laid out the way
 * Sources/Process/Cpu/Matrix_Mod/Create_Matrix.f90:87-119 builds it:
 *   - row(1) = 1, row(c+1) = row(c) + stencil width of cell c;
 *   - the diagonal is the FIRST entry of every row (dia(c) = row(c));
 *   - the remaining column indices of the row are sorted ascending;
 *   - int32 indices, FP64 values.
 * Cells of hexahedral grids are numbered lexicographically with x slowest and
 * z fastest, the order Sources/Shared/Grid_Mod/Sorting/Sort_Cells_By_Coordinates.f90
 * produces for the block-structured grids of Sources/Generate.
 *
 * Random numbers come from a counter-based generator (splitmix64 keyed by a
 * seed, a stream id and an index), so every input is reproducible bit for bit
 * on any machine and independent of the number of threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cases.h"

/* ------------------------------------------------------------------------ */
/*  Counter-based random numbers                                             */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Uniform double in [0,1) for (seed, stream, index). */
static inline double rng_u01(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t key = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull));
  return (double)(mix64(key + idx * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

double tfc_rng_u01(uint64_t seed, uint64_t stream, uint64_t idx) {
  return rng_u01(seed, stream, idx);
}

enum {
  ST_RHS = 1, ST_SPHERE = 2, ST_MODES = 3, ST_PECLET = 4, ST_JITTER = 5,
  ST_KNN_K = 6, ST_AREA = 7, ST_PERM = 8, ST_BATCH = 9
};

/* ------------------------------------------------------------------------ */
/*  Allocation helpers                                                       */
/* ------------------------------------------------------------------------ */
static int sys_alloc(tfc_system* s, int n, int64_t nnz) {
  memset(s, 0, sizeof(*s));
  if (n <= 0 || nnz <= 0 || nnz > 2147483000ll) return TFC_EINVAL;
  s->n = n;
  s->nnz = (int)nnz;
  s->row = (int*)malloc(sizeof(int) * ((size_t)n + 1));
  s->col = (int*)malloc(sizeof(int) * (size_t)nnz);
  s->dia = (int*)malloc(sizeof(int) * (size_t)n);
  s->val = (double*)malloc(sizeof(double) * (size_t)nnz);
  s->b = (double*)malloc(sizeof(double) * (size_t)n);
  if (!s->row || !s->col || !s->dia || !s->val || !s->b) {
    tfc_free(s);
    return TFC_ENOMEM;
  }
  return 0;
}

void tfc_free(tfc_system* s) {
  if (!s) return;
  free(s->row);
  free(s->col);
  free(s->dia);
  free(s->val);
  free(s->b);
  memset(s, 0, sizeof(*s));
}

/* U(-1,1) right-hand side; with mean_free the sequential mean is removed. */
static void fill_rhs(double* b, int n, uint64_t seed, int mean_free) {
  for (int i = 0; i < n; i++) b[i] = 2.0 * rng_u01(seed, ST_RHS, (uint64_t)i) - 1.0;
  if (mean_free) {
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += b[i];
    double mean = sum / (double)n;
    for (int i = 0; i < n; i++) b[i] -= mean;
  }
}

/* ------------------------------------------------------------------------ */
/*  7-point hexahedral structure                                             */
/* ------------------------------------------------------------------------ */
int64_t tfc_hex_nnz(int nx, int ny, int nz) {
  int64_t n = (int64_t)nx * ny * nz;
  return 7 * n - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

/* Fill row/col/dia for an nx*ny*nz grid; returns neighbour offsets order:
 * [diag, -x, -y, -z, +z, +y, +x] which is ascending column order because
 * the cell index is c = (ix*ny + iy)*nz + iz. */
static void hex_structure(int nx, int ny, int nz, tfc_system* s) {
  const int sy = nz, sx = ny * nz;
  int pos = 1; /* 1-based running position */
  int c = 0;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++, c++) {
        s->row[c] = pos;
        s->dia[c] = pos;
        s->col[pos - 1] = c + 1;
        pos++;
        if (ix > 0)      s->col[pos++ - 1] = c - sx + 1;
        if (iy > 0)      s->col[pos++ - 1] = c - sy + 1;
        if (iz > 0)      s->col[pos++ - 1] = c - 1 + 1;
        if (iz < nz - 1) s->col[pos++ - 1] = c + 1 + 1;
        if (iy < ny - 1) s->col[pos++ - 1] = c + sy + 1;
        if (ix < nx - 1) s->col[pos++ - 1] = c + sx + 1;
      }
  s->row[c] = pos;
}

/* Locate the CSR slot (0-based) of column cc (0-based) in row r (0-based). */
static inline int slot_of(const tfc_system* s, int r, int cc) {
  for (int j = s->row[r] - 1; j < s->row[r + 1] - 1; j++)
    if (s->col[j] == cc + 1) return j;
  return -1;
}

/*
 * Pressure-Poisson matrix on the unit cube, nx*ny*nz cells.
 * Face coefficient a_f = (A_f / d_f) * (1/rho)_f with (1/rho)_f the mean of
 * 1/rho of the two cells (harmonic mean of rho).  rho = 1 inside a sphere of
 * radius 0.25 whose centre is drawn U(0.25,0.75)^3 from the seed and
 * rho = rho_out outside (rho_out = 1 gives the constant-coefficient Laplacian).
 * Off-diagonals -a_f, diagonal sum of a_f: Neumann walls.  Cell 1 is pinned by
 * a Dirichlet face on its -x side: a_11 += A_f/(h_x/2) * (1/rho_1).
 * RHS U(-1,1) with its mean removed.
 */
int tfc_hex_pressure(int nx, int ny, int nz, double rho_out, uint64_t seed,
                     tfc_system* s) {
  if (nx < 1 || ny < 1 || nz < 1 || !(rho_out > 0.0)) return TFC_EINVAL;
  int64_t n64 = (int64_t)nx * ny * nz;
  if (n64 > 2000000000ll) return TFC_EINVAL;
  int n = (int)n64;
  int rc = sys_alloc(s, n, tfc_hex_nnz(nx, ny, nz));
  if (rc) return rc;
  hex_structure(nx, ny, nz, s);

  const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
  const double ax = hy * hz / hx, ay = hx * hz / hy, az = hx * hy / hz;
  double cx = 0.25 + 0.5 * rng_u01(seed, ST_SPHERE, 0);
  double cy = 0.25 + 0.5 * rng_u01(seed, ST_SPHERE, 1);
  double cz = 0.25 + 0.5 * rng_u01(seed, ST_SPHERE, 2);

  /* 1/rho per cell */
  double* irho = (double*)malloc(sizeof(double) * (size_t)n);
  if (!irho) { tfc_free(s); return TFC_ENOMEM; }
  int c = 0;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++, c++) {
        double x = (ix + 0.5) * hx - cx, y = (iy + 0.5) * hy - cy,
               z = (iz + 0.5) * hz - cz;
        int inside = (x * x + y * y + z * z) < 0.25 * 0.25;
        irho[c] = inside ? 1.0 : 1.0 / rho_out;
      }

  for (int64_t j = 0; j < s->nnz; j++) s->val[j] = 0.0;
  const int sy = nz, sx = ny * nz;
  c = 0;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++, c++) {
        /* faces towards higher neighbours, each visited once */
        int nb[3] = {iz < nz - 1 ? c + 1 : -1, iy < ny - 1 ? c + sy : -1,
                     ix < nx - 1 ? c + sx : -1};
        double af[3] = {az, ay, ax};
        for (int d = 0; d < 3; d++) {
          int c2 = nb[d];
          if (c2 < 0) continue;
          double a = af[d] * 0.5 * (irho[c] + irho[c2]);
          s->val[slot_of(s, c, c2)] -= a;
          s->val[slot_of(s, c2, c)] -= a;
          s->val[s->dia[c] - 1] += a;
          s->val[s->dia[c2] - 1] += a;
        }
      }
  /* Dirichlet pin of the first cell */
  s->val[s->dia[0] - 1] += 2.0 * ax * irho[0];
  free(irho);
  fill_rhs(s->b, n, seed, 1);
  return 0;
}

/* ------------------------------------------------------------------------ */
/*  Upwind convection-diffusion (non-symmetric momentum-like system)         */
/* ------------------------------------------------------------------------ */
typedef struct { double kx, ky, kz, wx, wy, wz, amp, phase; } mode_t_;

static void velocity(const mode_t_* m, int nm, double scale, double x, double y,
                     double z, double* u) {
  u[0] = u[1] = u[2] = 0.0;
  for (int k = 0; k < nm; k++) {
    double cph = cos(m[k].kx * x + m[k].ky * y + m[k].kz * z + m[k].phase);
    u[0] += m[k].amp * m[k].wx * cph;
    u[1] += m[k].amp * m[k].wy * cph;
    u[2] += m[k].amp * m[k].wz * cph;
  }
  u[0] *= scale; u[1] *= scale; u[2] *= scale;
}

/*
 * First-order upwind convection-diffusion on an N^3 grid of the unit cube.
 * Velocity: sum of 8 seeded Fourier modes amp_m * w_m * cos(k_m.x + phi_m)
 * with w_m perpendicular to k_m (so div u = 0 analytically) and the sum
 * scaled so |u| <= 1.  Face flux F = u(face centre).n * A_f, from the lower to
 * the higher cell.  Cell Peclet number Pe_c ~ logU(0.1, 50) sets the cell
 * viscosity nu_c = max(|u_c|,1e-3) h / Pe_c; face diffusion D_f = nu_f A_f / h
 * with nu_f the arithmetic mean.  Matrix (c1 < c2, F from c1 to c2):
 *   a(c1,c2) = -D_f - max(-F,0),  a(c2,c1) = -D_f - max(F,0),
 *   a(c1,c1) += D_f + max(F,0),   a(c2,c2) += D_f + max(-F,0),
 * plus the time term vol/dt on the diagonal with dt * max|u| / h = 0.5.
 * Walls carry neither flux nor diffusion.  RHS U(-1,1).
 */
int tfc_upwind(int N, uint64_t seed, tfc_system* s) {
  if (N < 2) return TFC_EINVAL;
  int64_t n64 = (int64_t)N * N * N;
  if (n64 > 2000000000ll) return TFC_EINVAL;
  int n = (int)n64;
  int rc = sys_alloc(s, n, tfc_hex_nnz(N, N, N));
  if (rc) return rc;
  hex_structure(N, N, N, s);
  const double h = 1.0 / N, two_pi = 6.283185307179586;

  mode_t_ m[8];
  double amp_sum = 0.0;
  for (int k = 0; k < 8; k++) {
    /* integer wave vector with components in {-2..2}, not all zero */
    int kv[3];
    for (int d = 0; d < 3; d++)
      kv[d] = (int)floor(5.0 * rng_u01(seed, ST_MODES, (uint64_t)(k * 16 + d))) - 2;
    if (kv[0] == 0 && kv[1] == 0 && kv[2] == 0) kv[2] = 1;
    double a[3];
    for (int d = 0; d < 3; d++) a[d] = 2.0 * rng_u01(seed, ST_MODES, (uint64_t)(1000 + k * 16 + d)) - 1.0;
    /* w = a x k, normalised */
    double wx = a[1] * kv[2] - a[2] * kv[1];
    double wy = a[2] * kv[0] - a[0] * kv[2];
    double wz = a[0] * kv[1] - a[1] * kv[0];
    double wn = sqrt(wx * wx + wy * wy + wz * wz);
    if (wn < 1e-12) { /* a parallel to k: pick any perpendicular */
      wx = -kv[1]; wy = kv[0]; wz = 0.0;
      if (kv[0] == 0 && kv[1] == 0) { wx = 1.0; wy = 0.0; }
      wn = sqrt(wx * wx + wy * wy + wz * wz);
    }
    m[k].kx = two_pi * kv[0]; m[k].ky = two_pi * kv[1]; m[k].kz = two_pi * kv[2];
    m[k].wx = wx / wn; m[k].wy = wy / wn; m[k].wz = wz / wn;
    m[k].amp = 0.5 + 0.5 * rng_u01(seed, ST_MODES, (uint64_t)(2000 + k));
    m[k].phase = two_pi * rng_u01(seed, ST_MODES, (uint64_t)(3000 + k));
    amp_sum += m[k].amp;
  }
  const double scale = 1.0 / amp_sum;

  double* nu = (double*)malloc(sizeof(double) * (size_t)n);
  if (!nu) { tfc_free(s); return TFC_ENOMEM; }
  double umax = 0.0;
  int c = 0;
  for (int ix = 0; ix < N; ix++)
    for (int iy = 0; iy < N; iy++)
      for (int iz = 0; iz < N; iz++, c++) {
        double u[3];
        velocity(m, 8, scale, (ix + 0.5) * h, (iy + 0.5) * h, (iz + 0.5) * h, u);
        double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (um > umax) umax = um;
        double lp = log(0.1) + (log(50.0) - log(0.1)) * rng_u01(seed, ST_PECLET, (uint64_t)c);
        nu[c] = (um > 1e-3 ? um : 1e-3) * h / exp(lp);
      }
  if (umax < 1e-12) umax = 1.0;
  const double dt = 0.5 * h / umax;
  const double area = h * h, vol = h * h * h;

  for (int64_t j = 0; j < s->nnz; j++) s->val[j] = 0.0;
  const int sy = N, sx = N * N;
  c = 0;
  for (int ix = 0; ix < N; ix++)
    for (int iy = 0; iy < N; iy++)
      for (int iz = 0; iz < N; iz++, c++) {
        s->val[s->dia[c] - 1] += vol / dt;
        int nb[3] = {iz < N - 1 ? c + 1 : -1, iy < N - 1 ? c + sy : -1,
                     ix < N - 1 ? c + sx : -1};
        for (int d = 0; d < 3; d++) {
          int c2 = nb[d];
          if (c2 < 0) continue;
          double fx = (ix + 0.5) * h, fy = (iy + 0.5) * h, fz = (iz + 0.5) * h;
          if (d == 0) fz += 0.5 * h;
          if (d == 1) fy += 0.5 * h;
          if (d == 2) fx += 0.5 * h;
          double u[3];
          velocity(m, 8, scale, fx, fy, fz, u);
          double F = (d == 0 ? u[2] : (d == 1 ? u[1] : u[0])) * area;
          double D = 0.5 * (nu[c] + nu[c2]) * area / h;
          double fpos = F > 0.0 ? F : 0.0, fneg = F < 0.0 ? -F : 0.0;
          s->val[slot_of(s, c, c2)] += -D - fneg;
          s->val[slot_of(s, c2, c)] += -D - fpos;
          s->val[s->dia[c] - 1] += D + fpos;
          s->val[s->dia[c2] - 1] += D + fneg;
        }
      }
  free(nu);
  fill_rhs(s->b, n, seed, 0);
  return 0;
}

/* ------------------------------------------------------------------------ */
/*  Unstructured kNN system                                                  */
/* ------------------------------------------------------------------------ */
static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/*
 * Jittered-lattice point cloud (nx*ny*nz points, spacing h = 1/nz, jitter
 * U(-0.3h,0.3h) per coordinate), each point connected to its k_c nearest
 * neighbours, k_c ~ U{kmin..kmax}, searched in the +-3 lattice window.  The
 * connectivity is symmetrised by union.  Face coefficient a_f = A_f / d_f,
 * A_f ~ logU(0.5,2) drawn per undirected pair, d_f the point distance.
 * Diagonal = sum of a_f; the first point (before permutation) is pinned by
 * a_11 += 2h.  Cells are then renumbered by a seeded random permutation, and
 * every row is rebuilt with the diagonal first and sorted columns.
 * RHS U(-1,1) with its mean removed.
 */
int tfc_knn(int nx, int ny, int nz, int kmin, int kmax, uint64_t seed,
            tfc_system* s) {
  memset(s, 0, sizeof(*s));
  if (nx < 4 || ny < 4 || nz < 4 || kmin < 1 || kmax < kmin || kmax > 64)
    return TFC_EINVAL;
  int64_t n64 = (int64_t)nx * ny * nz;
  if (n64 > 200000000ll) return TFC_EINVAL;
  const int n = (int)n64;
  const double h = 1.0 / nz;
  const int W = 3;

  double* px = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  int* kk = (int*)malloc(sizeof(int) * (size_t)n);
  int* nbr = (int*)malloc(sizeof(int) * (size_t)n * kmax);
  if (!px || !kk || !nbr) { free(px); free(kk); free(nbr); return TFC_ENOMEM; }

#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++) {
    int ix = c / (ny * nz), iy = (c / nz) % ny, iz = c % nz;
    px[3 * c + 0] = (ix + 0.5 + 0.6 * (rng_u01(seed, ST_JITTER, 3ull * c + 0) - 0.5)) * h;
    px[3 * c + 1] = (iy + 0.5 + 0.6 * (rng_u01(seed, ST_JITTER, 3ull * c + 1) - 0.5)) * h;
    px[3 * c + 2] = (iz + 0.5 + 0.6 * (rng_u01(seed, ST_JITTER, 3ull * c + 2) - 0.5)) * h;
    kk[c] = kmin + (int)floor((kmax - kmin + 1) * rng_u01(seed, ST_KNN_K, (uint64_t)c));
    if (kk[c] > kmax) kk[c] = kmax;
  }

#pragma omp parallel for schedule(dynamic, 1024)
  for (int c = 0; c < n; c++) {
    int ix = c / (ny * nz), iy = (c / nz) % ny, iz = c % nz;
    int k = kk[c];
    double bd[64];
    int bi[64];
    int cnt = 0;
    for (int dx = -W; dx <= W; dx++) {
      int jx = ix + dx;
      if (jx < 0 || jx >= nx) continue;
      for (int dy = -W; dy <= W; dy++) {
        int jy = iy + dy;
        if (jy < 0 || jy >= ny) continue;
        for (int dz = -W; dz <= W; dz++) {
          int jz = iz + dz;
          if (jz < 0 || jz >= nz) continue;
          int o = (jx * ny + jy) * nz + jz;
          if (o == c) continue;
          double ex = px[3 * o] - px[3 * c], ey = px[3 * o + 1] - px[3 * c + 1],
                 ez = px[3 * o + 2] - px[3 * c + 2];
          double d2 = ex * ex + ey * ey + ez * ez;
          /* insertion into the sorted top-k list; ties broken by index */
          if (cnt == k && (d2 > bd[k - 1] || (d2 == bd[k - 1] && o > bi[k - 1]))) continue;
          int p = cnt < k ? cnt++ : k - 1;
          while (p > 0 && (bd[p - 1] > d2 || (bd[p - 1] == d2 && bi[p - 1] > o))) {
            bd[p] = bd[p - 1];
            bi[p] = bi[p - 1];
            p--;
          }
          bd[p] = d2;
          bi[p] = o;
        }
      }
    }
    for (int j = 0; j < kmax; j++) nbr[(size_t)c * kmax + j] = j < cnt ? bi[j] : -1;
  }

  /* symmetric union: degree upper bound = own k + in-degree */
  int* deg = (int*)calloc((size_t)n + 1, sizeof(int));
  if (!deg) { free(px); free(kk); free(nbr); return TFC_ENOMEM; }
  for (int c = 0; c < n; c++)
    for (int j = 0; j < kmax; j++) {
      int o = nbr[(size_t)c * kmax + j];
      if (o < 0) continue;
      deg[c]++;
      deg[o]++;
    }
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!off) { free(px); free(kk); free(nbr); free(deg); return TFC_ENOMEM; }
  off[0] = 0;
  for (int c = 0; c < n; c++) off[c + 1] = off[c] + deg[c];
  int* adj = (int*)malloc(sizeof(int) * (size_t)off[n]);
  if (!adj) { free(px); free(kk); free(nbr); free(deg); free(off); return TFC_ENOMEM; }
  memset(deg, 0, sizeof(int) * (size_t)n);
  for (int c = 0; c < n; c++)
    for (int j = 0; j < kmax; j++) {
      int o = nbr[(size_t)c * kmax + j];
      if (o < 0) continue;
      adj[off[c] + deg[c]++] = o;
      adj[off[o] + deg[o]++] = c;
    }
  free(nbr);
  free(kk);
  /* sort and de-duplicate each adjacency list */
  int64_t nnz = n;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : nnz)
  for (int c = 0; c < n; c++) {
    int* a = adj + off[c];
    qsort(a, (size_t)deg[c], sizeof(int), cmp_int);
    int u = 0;
    for (int j = 0; j < deg[c]; j++)
      if (u == 0 || a[j] != a[u - 1]) a[u++] = a[j];
    deg[c] = u;
    nnz += u;
  }

  /* seeded random permutation: new index perm[c] */
  int* perm = (int*)malloc(sizeof(int) * (size_t)n);
  int* inv = (int*)malloc(sizeof(int) * (size_t)n);
  if (!perm || !inv) { free(px); free(deg); free(off); free(adj); free(perm); free(inv); return TFC_ENOMEM; }
  for (int c = 0; c < n; c++) inv[c] = c;
  for (int c = n - 1; c > 0; c--) {
    int j = (int)floor((c + 1) * rng_u01(seed, ST_PERM, (uint64_t)c));
    if (j > c) j = c;
    int t = inv[c]; inv[c] = inv[j]; inv[j] = t;
  }
  for (int c = 0; c < n; c++) perm[inv[c]] = c; /* inv[new] = old */

  int rc = sys_alloc(s, n, nnz);
  if (rc) { free(px); free(deg); free(off); free(adj); free(perm); free(inv); return rc; }
  s->row[0] = 1;
  for (int r = 0; r < n; r++) s->row[r + 1] = s->row[r] + 1 + deg[inv[r]];

#pragma omp parallel for schedule(dynamic, 4096)
  for (int r = 0; r < n; r++) {
    int c = inv[r];
    int base = s->row[r] - 1;
    int m = deg[c];
    int* cols = s->col + base + 1;
    for (int j = 0; j < m; j++) cols[j] = perm[adj[off[c] + j]];
    qsort(cols, (size_t)m, sizeof(int), cmp_int);
    double diag = 0.0;
    for (int j = 0; j < m; j++) {
      int o = inv[cols[j]];
      int lo = c < o ? c : o, hi = c < o ? o : c;
      double la = log(0.5) + (log(2.0) - log(0.5)) *
                  rng_u01(seed, ST_AREA, (uint64_t)lo * (uint64_t)n + (uint64_t)hi);
      double ex = px[3 * o] - px[3 * c], ey = px[3 * o + 1] - px[3 * c + 1],
             ez = px[3 * o + 2] - px[3 * c + 2];
      double a = exp(la) / sqrt(ex * ex + ey * ey + ez * ez);
      s->val[base + 1 + j] = -a;
      diag += a;
      cols[j] += 1; /* 1-based */
    }
    if (c == 0) diag += 2.0 * h;
    s->col[base] = r + 1;
    s->dia[r] = base + 1;
    s->val[base] = diag;
  }
  free(px); free(deg); free(off); free(adj); free(perm); free(inv);
  fill_rhs(s->b, n, seed, 1);
  return 0;
}

/* ------------------------------------------------------------------------ */
/*  Batch of pressure systems                                                */
/* ------------------------------------------------------------------------ */
/* Side lengths s_k ~ U{smin..smax} and density ratios logU(1, rmax). */
void tfc_batch_params(int count, int smin, int smax, double rmax, uint64_t seed,
                      int* sides, double* ratios) {
  for (int k = 0; k < count; k++) {
    int sd = smin + (int)floor((smax - smin + 1) * rng_u01(seed, ST_BATCH, 2ull * k));
    sides[k] = sd > smax ? smax : sd;
    ratios[k] = exp(log(rmax) * rng_u01(seed, ST_BATCH, 2ull * k + 1));
  }
}
