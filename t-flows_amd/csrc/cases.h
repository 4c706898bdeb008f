/*
 * Synthetic linear systems in T-Flows' 1-based CSR layout
 * (Sources/Process/Cpu/Matrix_Mod/Create_Matrix.f90:87-119: diagonal first,
 * remaining columns ascending).  Host-side test and benchmark inputs.
 */
#ifndef TFLOWS_CASES_H
#define TFLOWS_CASES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TFC_EINVAL = -1, TFC_ENOMEM = -2 };

/* A generated system.  All arrays are malloc'd by the generator and released
 * by tfc_free().  row: n+1, col: nnz, dia: n (positions), val: nnz, b: n. */
typedef struct {
  int n;
  int nnz;
  int* row;
  int* col;
  int* dia;
  double* val;
  double* b;
} tfc_system;

double tfc_rng_u01(uint64_t seed, uint64_t stream, uint64_t idx);
int64_t tfc_hex_nnz(int nx, int ny, int nz);
int tfc_hex_pressure(int nx, int ny, int nz, double rho_out, uint64_t seed, tfc_system* s);
int tfc_upwind(int N, uint64_t seed, tfc_system* s);
int tfc_knn(int nx, int ny, int nz, int kmin, int kmax, uint64_t seed, tfc_system* s);
void tfc_batch_params(int count, int smin, int smax, double rmax, uint64_t seed,
                      int* sides, double* ratios);
void tfc_free(tfc_system* s);

#ifdef __cplusplus
}
#endif
#endif
