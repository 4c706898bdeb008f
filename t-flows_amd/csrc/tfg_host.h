// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <string>
#include <vector>

#include "tfg_internal.h"

namespace tfg {

// Records the message returned by tfg_last_error() (per thread); returns code.
int fail(int code, const char* fmt, ...);

// Fortran CHARACTER comparison ignores trailing blanks.
std::string fstring(const char* s, int len);

// Greedy row blocks: at most 256 rows and at most cap nonzeros each; a single
// longer row forms a block of its own.  row1: 1-based row starts (only
// differences matter); blk receives the block boundaries, 0-based rows.
void make_blocks(const int* row1, int n, int cap, std::vector<int>& blk);

// The layout Matrix_Type guarantees (Create_Matrix.f90:87-119): 1-based,
// rows non-empty, diagonal first, columns in range.
int check_csr(int nt, int nnz, const int* row, const int* col, const int* dia);

}  // namespace tfg

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return tfg::fail(e_ == hipErrorOutOfMemory ? TFG_ENOMEM : TFG_EHIP, "%s: %s (%s:%d)", \
                       #expr, hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)
