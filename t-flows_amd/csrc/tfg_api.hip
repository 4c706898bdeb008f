// C-ABI of tflows_gpu: structure upload, the Solve_Native drop-in and the
// device building blocks.  Host orchestration only; kernels live in
// tfg_kernels.hip (FP64) and tfg_ir.hip (FP32 inner / FP64 refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <utility>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "tfg_host.h"

using tfg::State;
using tfg::fail;
using tfg::fstring;
using tfg::make_blocks;

namespace tfg {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

std::string fstring(const char* s, int len) {
  if (!s) return std::string();
  size_t n = len < 0 ? strlen(s) : (size_t)len;
  while (n > 0 && (s[n - 1] == ' ' || s[n - 1] == '\0')) --n;
  return std::string(s, n);
}

void make_blocks(const int* row1, int n, int cap, std::vector<int>& blk) {
  blk.clear();
  blk.push_back(0);
  int r = 0;
  while (r < n) {
    int start = r;
    long long s = row1[r];
    while (r < n && r - start < kRowsPerBlock && (long long)row1[r + 1] - s <= cap) ++r;
    if (r == start) ++r;  // long row
    blk.push_back(r);
  }
}

int check_csr(int nt, int nnz, const int* row, const int* col, const int* dia) {
  if (row[0] != 1) return fail(TFG_EINVAL, "row(1) must be 1");
  if ((long long)row[nt] - 1 != nnz || nnz < nt)
    return fail(TFG_EINVAL, "nnz=%d does not match row(nt+1)-1=%d", nnz, row[nt] - 1);
  for (int i = 0; i < nt; ++i) {
    if (row[i + 1] <= row[i]) return fail(TFG_EINVAL, "row %d is empty or row not increasing", i + 1);
    if (dia[i] != row[i] || col[row[i] - 1] != i + 1)
      return fail(TFG_EINVAL, "diagonal must be the first entry of row %d", i + 1);
  }
  for (long long j = 0; j < nnz; ++j)
    if (col[j] < 1 || col[j] > nt) return fail(TFG_EINVAL, "col(%lld)=%d out of range", j + 1, col[j]);
  return TFG_OK;
}

}  // namespace tfg

// Scratch of the allocation-free solve (tfg_solve_native_noalloc): one
// static device array, of which every GPU holds its own copy.  In host code
// the array's name is a host-side stand-in, not a device address, so it is
// never passed to a kernel or a copy: set-up resolves each device's copy
// once, with that device current, and only that address is used.
static __device__ __align__(256) unsigned char scratch[TFG_NOALLOC_MAX_BYTES];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kScratchDevices] = {};

// Device address of `device`'s scratch array, resolved on its first call
// (tfg_matrix_create, with `device` current).
int scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail(TFG_EINVAL, "device %d: the scratch table holds %d devices", device,
                kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return TFG_OK;
  void* ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(TFG_EHIP, "no device address for the scratch array on device %d", device);
  g_scratch[device] = static_cast<unsigned char*>(ptr);
  return TFG_OK;
}

unsigned char* scratch_of(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

}  // namespace

// What a captured chunk of CG iterations depends on besides the handle.
struct GraphKey {
  const void* val;
  const void* b;
  int prec, dx, chunk, kind, sym;
  bool operator==(const GraphKey& o) const {
    return val == o.val && b == o.b && prec == o.prec && dx == o.dx && chunk == o.chunk &&
           kind == o.kind && sym == o.sym;
  }
};

struct tfg_matrix {
  int device = 0;
  int ni = 0, nt = 0, nnz = 0;
  int cu_count = 0;
  hipStream_t stream = nullptr;
  // structure (0-based)
  int* row = nullptr;
  int* col = nullptr;
  int* blk[2] = {nullptr, nullptr};  // stage row blocks of kCap / kCap2 nonzeros
  int nblk[2] = {0, 0};
  int grid[tfg::kKinds] = {};         // resident SpMV workgroups per engine
  int kind = -1;                      // chosen SpMV engine (-1: tune on first use)
  float tune_us[tfg::kKinds] = {};
  // FP64 work
  double* val = nullptr;  // host-path copy of val
  double* b = nullptr;    // host-path copy of b
  double* x = nullptr;
  double* r = nullptr;
  double* p = nullptr;
  double* q = nullptr;
  double* ad = nullptr;
  double* part = nullptr;     // partial sums of every phase (see Partials)
  int pstride = 0;
  State* st = nullptr;       // device scalars of the solve
  State* st_in = nullptr;    // device scalars of an inner (FP32) solve
  State* hst = nullptr;      // pinned: [0] set-up, [1],[2] polling, [3] final, [4] inner
  hipEvent_t ev_poll[2] = {nullptr, nullptr};
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
  std::vector<hipEvent_t> ev_sample;
  // FP32 sweeps of the refinement form (allocated on first use): the FP32
  // matrix (CSR val32, or the symmetric storage sd32 with its 4-row blocks),
  // row sums, 1 / a_ii, vectors
  float* val32 = nullptr;
  float* sd32 = nullptr;
  long long stride32 = 0;
  int* sorder32 = nullptr;
  int nsorder32 = 0;
  int quad32 = 0;
  int grid32 = 0;
  double* s64 = nullptr;
  float* s32 = nullptr;
  float* e32 = nullptr;
  float* r32 = nullptr;
  float* p32 = nullptr;
  float* q32 = nullptr;
  float* ad32 = nullptr;
  float* ns32[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // textbook methods
  float* valt32 = nullptr;                                         // BiCG: A^T
  // textbook BiCG / CGS / BiCGSTAB work (allocated on first use)
  double* rt = nullptr;
  double* nv[4] = {nullptr, nullptr, nullptr, nullptr};
  double* valt = nullptr;
  int* tpos = nullptr;
  int tpos_state = 0;  // 0 not built, 1 built, -1 the pattern is not symmetric
  // DIC levels (built at create) and work
  int* icc_rows[2] = {nullptr, nullptr};   // forward, backward
  std::vector<int> icc_off[2];
  double* icc_d = nullptr;
  double* icc_dinv = nullptr;
  double* icc_y = nullptr;
  double* icc_z = nullptr;
  bool box = false;               // box grid: tiled DIC sweeps (bx, box_toff)
  tfg::BoxIcc bx{};
  std::vector<int> box_toff;
  // graph path: executable chunk of CG iterations and the operands it holds
  hipGraphExec_t cg_graph = nullptr;
  GraphKey cg_key{};
  // multi-rank hooks
  tfg_allreduce_fn allreduce = nullptr;
  tfg_exchange_fn exchange = nullptr;
  void* comm_ctx = nullptr;
  double* loc = nullptr;   // device: this rank's totals of a reduction
  double* vmax = nullptr;  // device: tfg_vel_max's partials and result
  double* vrange = nullptr;  // device: tfg_range's partials and result
  double* icc_r64 = nullptr;  // FP32 form with ICC: the widened FP32 residual
  float* z32 = nullptr;       // FP32 form with ICC: z = M^-1 r rounded to FP32
  // offset-coded columns (structured grids): col = row + tab[code]
  uint8_t* code = nullptr;
  int* tab = nullptr;
  int ntab = 0;        // 0: more than 256 distinct offsets, no coded engine
  // row patterns (structured grids): each row's offset sequence by id
  uint8_t* pat = nullptr;
  int* poff = nullptr;
  int* plen = nullptr;
  int npat = 0;        // 0: no pattern engine on this matrix
  // symmetric engine: slot of each pattern entry, diagonal + upper diagonals
  // (allocated on first use), asymmetry flag and its pinned host copy
  int* pslot = nullptr;
  int* sorder = nullptr;  // (start, end) row blocks in processing order
  int nsorder = 0;
  tfg::SymDiag sdg{};     // symmetric engine geometry (offsets, stride, mode)
  int nsym = 0;        // upper diagonals; 0: no symmetric engine
  double* sym = nullptr;
  int* sym_flag = nullptr;
  int* hflag = nullptr;
  int sym_state = -1;  // last val: 1 symmetric, 0 not, -1 none checked yet
  // internal ordering (tfg_matrix_reorder): the handle of the renumbered
  // matrix, new <- old maps of values (vmap) and rows (iperm), old <- new of
  // rows (perm), and permuted copies of val, b, x and an output vector
  tfg_matrix* inner = nullptr;
  int *vmap = nullptr, *perm = nullptr, *iperm = nullptr;
  double *pval = nullptr, *pb = nullptr, *px = nullptr, *pq = nullptr;
};

namespace {

// Rows of the partials buffer: diag 1, prologue 7, spmv 1, update 5, aux 2,
// ext 6, irc 4.
constexpr int kPartRows = 26;

// The partial rows in the handle's buffer, or in `base` (same layout).
tfg::Partials partials_of(const tfg_matrix* h, const tfg::Launch& L, double* base = nullptr) {
  tfg::Partials pt;
  pt.pstride = h->pstride;
  pt.diag = base ? base : h->part;
  pt.pro = pt.diag + 1 * (size_t)h->pstride;
  pt.spmv = pt.pro + 7 * (size_t)h->pstride;
  pt.upd = pt.spmv + 1 * (size_t)h->pstride;
  pt.aux = pt.upd + 5 * (size_t)h->pstride;
  pt.ext = pt.aux + 2 * (size_t)h->pstride;
  pt.irc = pt.ext + 6 * (size_t)h->pstride;
  pt.g_vec = L.grid_vec;
  pt.g_spmv = L.grid_spmv;
  pt.g_upd = L.grid_vec;
  return pt;
}


// Rows a workgroup of each engine covers per pass (stage: one block of up to
// 256 rows; wave: 256 / L rows).
int rows_per_pass(int kind) {
  switch (kind) {
    case tfg::K_WAVE4: return tfg::kBlock / 4;
    case tfg::K_WAVE8: return tfg::kBlock / 8;
    case tfg::K_WAVE16: return tfg::kBlock / 16;
    default: return tfg::kRowsPerBlock;
  }
}

void size_grids(tfg_matrix* h) {
  for (int k = 0; k < tfg::kKinds; ++k) {
    const int per_cu = tfg::spmv_blocks_per_cu(k);
    const int work = (k == tfg::K_STAGE2048 || k == tfg::K_PIPE2048 || k == tfg::K_PIPE2048_CODED ||
                      k == tfg::K_PIPE_PATTERN)
                         ? h->nblk[0]
                     : k == tfg::K_SYM_DIA ? std::max(1, h->nsorder)
                     : k == tfg::K_STAGE4096 ? h->nblk[1]
                                             : (h->nt + rows_per_pass(k) - 1) / rows_per_pass(k);
    int gs = std::min(work, std::max(1, h->cu_count) * per_cu);
    if (gs >= tfg::kXcds) gs -= gs % tfg::kXcds;
    h->grid[k] = std::max(1, gs);
  }
}

// Grid of the element-wise kernels: at most 3 workgroups per CU.  Every
// workgroup of the next phase sweeps all of their partials (5 per
// workgroup), so more workgroups cost more than they hide: at 200^3 the
// solve took 0.469 / 0.468 / 0.473 / 0.494 / 0.545 s with 2 / 3 / 4 / 8 /
// 16 per CU.
int grid_vec_of(const tfg_matrix* h) {
  const int pairs = (h->nt + 1) / 2;
  const int gv = (pairs + tfg::kBlock - 1) / tfg::kBlock;
  return std::max(1, std::min(gv, std::max(1, h->cu_count) * 3));
}

tfg::Launch launch_of(const tfg_matrix* h, const void* val, int kind) {
  tfg::Launch L;
  L.stream = h->stream;
  L.kind = kind;
  L.grid_spmv = h->grid[kind];
  L.grid_vec = grid_vec_of(h);
  L.vec_val = ((uintptr_t)val % 16) == 0 && h->nnz >= 4;
  L.sym = false;
  return L;
}

tfg::SpmvShape shape_of(const tfg_matrix* h, int kind) {
  tfg::SpmvShape m;
  m.n = h->ni;
  m.row = h->row;
  m.col = h->col;
  const int b = kind == tfg::K_STAGE4096 ? 1 : 0;
  m.blk = h->blk[b];
  m.nblk = h->nblk[b];
  m.code = h->code;
  m.tab = h->tab;
  m.ntab = h->ntab;
  m.pat = h->pat;
  m.poff = h->poff;
  m.plen = h->plen;
  m.pslot = h->pslot;
  m.sd = h->sym;
  m.sorder = h->sorder;
  m.nsorder = h->nsorder;
  m.sdg = h->sdg;
  return m;
}

// CgArgs with the structure of the chosen engine filled in.
tfg::CgArgs args_of(const tfg_matrix* h, int kind) {
  const tfg::SpmvShape m = shape_of(h, kind);
  tfg::CgArgs a{};
  a.n = h->ni;
  a.row = h->row;
  a.col = h->col;
  a.blk = m.blk;
  a.nblk = m.nblk;
  a.code = m.code;
  a.tab = m.tab;
  a.ntab = m.ntab;
  a.pat = m.pat;
  a.poff = m.poff;
  a.plen = m.plen;
  a.pslot = m.pslot;
  a.sd = m.sd;
  a.sorder = m.sorder;
  a.nsorder = m.nsorder;
  a.sdg = m.sdg;
  return a;
}

// Engines this matrix offers (6 needs offset codes, 7 row patterns).
bool kind_available(const tfg_matrix* h, int k) {
  if (k == tfg::K_PIPE2048_CODED) return h->ntab > 0;
  if (k == tfg::K_PIPE_PATTERN) return h->npat > 0;
  if (k == tfg::K_SYM_DIA) return h->nsym > 0;
  return k >= 0 && k < tfg::kKinds;
}

int alloc_work(tfg_matrix* h) {
  const size_t n = (size_t)h->nt;
  HIP_TRY(hipMalloc(&h->x, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->r, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->p, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->q, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->ad, n * sizeof(double)));
  size_grids(h);
  h->pstride = grid_vec_of(h);
  for (int k = 0; k < tfg::kKinds; ++k) h->pstride = std::max(h->pstride, h->grid[k]);
  HIP_TRY(hipMalloc(&h->part, sizeof(double) * kPartRows * (size_t)h->pstride));
  HIP_TRY(hipMemset(h->part, 0, sizeof(double) * kPartRows * (size_t)h->pstride));
  HIP_TRY(hipMalloc(&h->st, sizeof(State)));
  HIP_TRY(hipMalloc(&h->st_in, sizeof(State)));
  HIP_TRY(hipMemset(h->st, 0, sizeof(State)));
  HIP_TRY(hipMemset(h->st_in, 0, sizeof(State)));
  HIP_TRY(hipHostMalloc(&h->hst, 6 * sizeof(State), hipHostMallocDefault));
  memset(h->hst, 0, 6 * sizeof(State));
  HIP_TRY(hipEventCreate(&h->ev_poll[0]));
  HIP_TRY(hipEventCreate(&h->ev_poll[1]));
  HIP_TRY(hipEventCreate(&h->ev_t0));
  HIP_TRY(hipEventCreate(&h->ev_t1));
  return TFG_OK;
}

// Teardown in dependency order: wait for the handle's work, then the
// executable graph (its kernel nodes hold this handle's buffers as
// arguments) and the inner handle, then the memory, events and stream.
void free_all(tfg_matrix* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->cg_graph) (void)hipGraphExecDestroy(h->cg_graph);
  h->cg_graph = nullptr;
  if (h->inner) free_all(h->inner);
  h->inner = nullptr;
  void* dev[] = {h->row, h->col, h->blk[0], h->blk[1], h->val, h->b, h->x, h->r, h->p, h->q, h->ad,
                 h->part, h->st, h->st_in, h->val32, h->sd32, h->sorder32, h->s64, h->s32, h->e32,
                 h->r32, h->p32, h->q32, h->ad32, h->ns32[0], h->ns32[1], h->ns32[2],
                 h->ns32[3], h->ns32[4], h->valt32, h->rt, h->nv[0], h->nv[1], h->nv[2],
                 h->nv[3], h->valt, h->tpos, h->code, h->tab, h->pat, h->poff, h->plen,
                 h->pslot, h->sym, h->sym_flag, h->sorder,
                 h->loc, h->vmax, h->vrange, h->icc_r64, h->z32, h->icc_rows[0],
                 h->icc_rows[1], h->icc_d, h->icc_dinv, h->icc_y, h->icc_z,
                 (void*)h->bx.tiles, h->bx.dg, h->bx.lo[0], h->bx.lo[1],
                 h->bx.lo[2], h->bx.up[0], h->bx.up[1], h->bx.up[2]};
  for (void* d : dev)
    if (d) (void)hipFree(d);
  if (h->hst) (void)hipHostFree(h->hst);
  if (h->hflag) (void)hipHostFree(h->hflag);
  for (void* d : {(void*)h->vmap, (void*)h->perm, (void*)h->iperm, (void*)h->pval,
                  (void*)h->pb, (void*)h->px, (void*)h->pq})
    if (d) (void)hipFree(d);
  for (hipEvent_t e : {h->ev_poll[0], h->ev_poll[1], h->ev_t0, h->ev_t1})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->ev_sample)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  (void)hipGetLastError();  // release errors are dropped, not left for later calls
  delete h;
}

// Engine choice.  By default the pipelined ~2K-nonzero stream engine (the
// fastest on every T-Flows-like structure measured).  The choice fixes the
// SpMV launch grid and with it the tree of the fused dot products, so a
// timed choice could differ between two runs on the same matrix and move
// the last bits of the result; timing is therefore opt-in: tfg_spmv_tune(),
// or TFG_SPMV_AUTOTUNE=1 in the environment (times every engine, 2 warm-up +
// 5 timed launches each, keeps the fastest).  TFG_SPMV_KIND=<0..5> forces one.
int tune_all(tfg_matrix* h, const double* val);

// The engine chosen without timing: the symmetric one at 2M nonzeros and
// more, else the most compact format the matrix offers.
int default_kind(const tfg_matrix* h) {
  return h->nsym > 0 && h->nnz >= (1 << 21) ? tfg::K_SYM_DIA
         : h->npat > 0                      ? tfg::K_PIPE_PATTERN
         : h->ntab > 0                      ? tfg::K_PIPE2048_CODED
                                            : tfg::K_PIPE2048;
}

int ensure_tuned(tfg_matrix* h, const double* val) {
  if (h->kind >= 0) return TFG_OK;
  const char* at = getenv("TFG_SPMV_AUTOTUNE");
  if (!(at && *at && *at != '0')) {
    h->kind = default_kind(h);
    return TFG_OK;
  }
  return tune_all(h, val);
}

// Symmetric engine for one val: fills the diagonal and upper diagonals and
// reports whether val is symmetric bit for bit (one host wait on a flag).
// *ok = false: the products must use the row-pattern engine.
int sym_prepare(tfg_matrix* h, const double* val, bool* ok) {
  *ok = false;
  if (h->nsym <= 0) return TFG_OK;
  if (!h->sym) {
    // zero once: entries a row lacks and the gaps between slots stay zero
    const size_t bytes = sizeof(double) * (size_t)(h->nsym + 1) * (size_t)h->sdg.stride;
    HIP_TRY(hipMalloc(&h->sym, bytes));
    HIP_TRY(hipMemsetAsync(h->sym, 0, bytes, h->stream));
    HIP_TRY(hipMalloc(&h->sym_flag, sizeof(int)));
    HIP_TRY(hipHostMalloc(&h->hflag, sizeof(int), hipHostMallocDefault));
  }
  HIP_TRY(hipMemsetAsync(h->sym_flag, 0, sizeof(int), h->stream));
  HIP_TRY(tfg::launch_sym_pack(h->stream, grid_vec_of(h), shape_of(h, tfg::K_SYM_DIA), val, h->sym,
                               h->sym_flag));
  HIP_TRY(hipMemcpyAsync(h->hflag, h->sym_flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *ok = *h->hflag == 0;
  h->sym_state = *ok ? 1 : 0;
  return TFG_OK;
}

int tune_all(tfg_matrix* h, const double* val) {
  HIP_TRY(hipMemsetAsync(h->p, 0, sizeof(double) * (size_t)h->nt, h->stream));
  float best = 1e30f;
  int best_k = tfg::K_STAGE2048;
  for (int k = 0; k < tfg::kKinds; ++k) {
    if (!kind_available(h, k)) {
      h->tune_us[k] = -1.f;  // not available on this matrix
      continue;
    }
    tfg::Launch L = launch_of(h, val, k);
    if (k == tfg::K_SYM_DIA) {
      bool ok = false;
      const int rs = sym_prepare(h, val, &ok);
      if (rs != TFG_OK) return rs;
      if (!ok) {
        h->tune_us[k] = -1.f;  // val is not symmetric
        continue;
      }
      L.sym = true;
    }
    const tfg::SpmvShape m = shape_of(h, k);
    for (int w = 0; w < 2; ++w) HIP_TRY(tfg::launch_spmv_plain(L, m, val, h->p, h->q));
    HIP_TRY(hipEventRecord(h->ev_t0, h->stream));
    for (int w = 0; w < 5; ++w) HIP_TRY(tfg::launch_spmv_plain(L, m, val, h->p, h->q));
    HIP_TRY(hipEventRecord(h->ev_t1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev_t1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
    h->tune_us[k] = ms * 1000.f / 5.f;
    if (h->tune_us[k] < best) {
      best = h->tune_us[k];
      best_k = k;
    }
  }
  h->kind = best_k;
  return TFG_OK;
}

// Position of every entry's transposed entry (built once per handle on the
// device); h->tpos_state -1 when the pattern is not symmetric.
int ensure_tpos(tfg_matrix* h) {
  if (h->tpos_state == 0) {
    HIP_TRY(hipMalloc(&h->tpos, sizeof(int) * (size_t)h->nnz));
    int* bad = nullptr;
    HIP_TRY(hipMalloc(&bad, sizeof(int)));
    int hb = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), h->stream);
    if (e == hipSuccess) e = tfg::launch_transpose_pos(h->stream, h->nt, h->row, h->col, h->tpos, bad);
    if (e == hipSuccess) e = hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(bad);
    HIP_TRY(e);
    h->tpos_state = hb ? -1 : 1;
  }
  return TFG_OK;
}

// Work vectors of the textbook methods; for BiCG also the transposed-entry
// map and the transposed values.
int ensure_nonsym(tfg_matrix* h, int method, const double* val, const tfg::Launch& L) {
  const size_t n = (size_t)h->nt;
  if (!h->rt) HIP_TRY(hipMalloc(&h->rt, n * sizeof(double)));
  for (double*& v : h->nv)
    if (!v) HIP_TRY(hipMalloc(&v, n * sizeof(double)));
  if (method != tfg::NS_BICG) return TFG_OK;
  { const int rt = ensure_tpos(h); if (rt != TFG_OK) return rt; }
  if (h->tpos_state < 0)
    return fail(TFG_ENOTSUP, "textbook BiCG needs a structurally symmetric matrix");
  if (!h->valt) HIP_TRY(hipMalloc(&h->valt, sizeof(double) * (size_t)h->nnz));
  HIP_TRY(tfg::launch_gather_valt(L, h->nnz, h->tpos, val, h->valt));
  return TFG_OK;
}

// Dependency levels of the DIC sweeps, read exactly as the sweeps read rows
// (Prec_Solve.f90:64-92): forward, row i needs the entries after its
// diagonal up to the first column above i; backward, the entries from the
// row's end down to the first column below i.  row0/col0: 0-based host.
void icc_levels(const int* row0, const int* col0, int n, bool forward, std::vector<int>& rows,
                std::vector<int>& off) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  for (int t = 0; t < n; ++t) {
    const int i = forward ? t : n - 1 - t;
    int l = 0;
    if (forward) {
      for (int j = row0[i] + 1; j < row0[i + 1]; ++j) {
        const int k = col0[j];
        if (k > i) break;
        l = std::max(l, lev[k] + 1);
      }
    } else {
      for (int j = row0[i + 1] - 1; j > row0[i]; --j) {
        const int k = col0[j];
        if (k < i) break;
        if (k < n) l = std::max(l, lev[k] + 1);
      }
    }
    lev[i] = l;
    nlev = std::max(nlev, l + 1);
  }
  off.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i) off[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) off[l + 1] += off[l];
  std::vector<int> pos(off.begin(), off.end() - 1);
  rows.assign(n, 0);
  for (int i = 0; i < n; ++i) rows[pos[lev[i]]++] = i;
}

int ensure_icc(tfg_matrix* h) {
  if (!h->icc_rows[0] || !h->icc_rows[1])
    return fail(TFG_ENOTSUP, "no DIC levels on this handle");
  const size_t n = (size_t)h->nt;
  for (double** v : {&h->icc_d, &h->icc_dinv, &h->icc_y, &h->icc_z})
    if (!*v) HIP_TRY(hipMalloc(v, n * sizeof(double)));
  if (h->box && !h->bx.dg) {  // zero once: entries a row lacks stay zero
    for (double** v : {&h->bx.dg, &h->bx.lo[0], &h->bx.lo[1], &h->bx.lo[2], &h->bx.up[0],
                       &h->bx.up[1], &h->bx.up[2]}) {
      HIP_TRY(hipMalloc(v, n * sizeof(double)));
      HIP_TRY(hipMemsetAsync(*v, 0, n * sizeof(double), h->stream));
    }
  }
  return TFG_OK;
}



tfg::IccArgs icc_args(const tfg_matrix* h) {
  tfg::IccArgs c{};
  c.n = h->ni;
  c.row = h->row;
  c.col = h->col;
  c.rows_f = h->icc_rows[0];
  c.rows_b = h->icc_rows[1];
  c.d = h->icc_d;
  c.dinv = h->icc_dinv;
  c.y = h->icc_y;
  c.nlev_f = (int)h->icc_off[0].size() - 1;
  c.nlev_b = (int)h->icc_off[1].size() - 1;
  c.off_f_host = h->icc_off[0].data();
  c.off_b_host = h->icc_off[1].data();
  c.box = h->box ? 1 : 0;
  c.bx = h->bx;
  c.toff_host = h->box_toff.data();
  return c;
}

// Processing order of the symmetric engines: planes of D rows (D = the
// largest offset; plane starts rounded down to a multiple of `align` rows),
// blocks of at most R rows inside a plane, taken band by band (8
// consecutive blocks of a plane) and, within a band, plane after plane.
// The workgroups that run at the same time then cover one band over
// neighbouring planes, so most rows a row reads (one plane or one line
// away) are read by them too and stay in L2: 377 instead of 522 MB fetched
// per FP64 product at 200^3.  (start, end) pairs; every row exactly once,
// starts and lengths multiples of align (ni a multiple of align); empty if
// that fails.
std::vector<int> sym_blocks(int ni, int D, int R, int align) {
  // a plane may gain up to align - 1 rows (at least one is allowed for)
  const long long bpp = ((long long)D + std::max(align, 2) - 1 + R - 1) / R;
  constexpr int kBand = 8;
  const long long amask = ~((long long)align - 1);
  std::vector<int> ord;
  if (D >= kBand * R) {
    const long long planes = ((long long)ni + D - 1) / D;
    auto start = [&](long long x) { return std::min<long long>(ni, (x * D) & amask); };
    for (long long j0 = 0; j0 < bpp; j0 += kBand)
      for (long long x = 0; x < planes; ++x)
        for (long long b = j0; b < std::min<long long>(bpp, j0 + kBand); ++b) {
          const long long r0 = start(x) + b * R;
          const long long r1 = std::min<long long>({r0 + R, start(x + 1), (long long)ni});
          if (r0 < r1) {
            ord.push_back((int)r0);
            ord.push_back((int)r1);
          }
        }
  } else {
    for (int r0 = 0; r0 < ni; r0 += R) {
      ord.push_back(r0);
      ord.push_back(std::min(ni, r0 + R));
    }
  }
  std::vector<char> seen((size_t)ni, 0);
  for (size_t t = 0; t + 1 < ord.size(); t += 2) {
    if ((ord[t] % align) || ((ord[t + 1] - ord[t]) % align)) return {};
    for (int r = ord[t]; r < ord[t + 1]; ++r) {
      if (seen[r]) return {};
      seen[r] = 1;
    }
  }
  for (int r = 0; r < ni; ++r)
    if (!seen[r]) return {};
  return ord;
}

// Box grid of the tiled DIC sweeps: rows (x, y, z) of an nx*ny*nz grid, x
// slowest, every row stored diagonal first and then by ascending column,
// every off-diagonal entry a box neighbour (offsets 1, nz, ny*nz, inside the
// box), no buffer cells.  The sides come from the distinct positive offsets
// d0 = 1 < d1 < d2: nz = d1, ny = d2 / d1 (fewer offsets: a 2-D or 1-D
// grid).  Sets h->box and the tile tables; false leaves the row-level
// sweeps.
int box_setup(tfg_matrix* h, const int* row0, const int* col0) {
  const int n = h->ni;
  if (n != h->nt || n < 2) return TFG_OK;
  std::vector<long long> offs;
  for (int i = 0; i < n; ++i) {
    if (row0[i + 1] <= row0[i] || col0[row0[i]] != i) return TFG_OK;  // diagonal first
    for (int j = row0[i] + 1; j < row0[i + 1]; ++j) {
      if (j > row0[i] + 1 && !(col0[j - 1] < col0[j])) return TFG_OK;  // ascending
      const long long d = (long long)col0[j] - i;
      if (d > 0 && std::find(offs.begin(), offs.end(), d) == offs.end()) {
        if (offs.size() == 3) return TFG_OK;
        offs.push_back(d);
      }
    }
  }
  std::sort(offs.begin(), offs.end());
  if (offs.empty() || offs[0] != 1) return TFG_OK;
  long long nz = n, ny = 1, nx = 1;
  if (offs.size() >= 2) {
    nz = offs[1];
    if (n % nz) return TFG_OK;
    ny = n / nz;
  }
  if (offs.size() == 3) {
    if (offs[2] % offs[1]) return TFG_OK;
    ny = offs[2] / offs[1];
    if (n % offs[2]) return TFG_OK;
    nx = n / offs[2];
  }
  const long long nyz = ny * nz;
  for (int i = 0; i < n; ++i) {
    const long long x = i / nyz, y = (i / nz) % ny, z = i % nz;
    for (int j = row0[i] + 1; j < row0[i + 1]; ++j) {
      const long long d = (long long)col0[j] - i;
      const bool ok = (d == 1 && z + 1 < nz) || (d == -1 && z > 0) || (d == nz && y + 1 < ny) ||
                      (d == -nz && y > 0) || (d == nyz && x + 1 < nx) || (d == -nyz && x > 0);
      if (!ok) return TFG_OK;
    }
  }
  tfg::BoxIcc& b = h->bx;
  b = tfg::BoxIcc{};
  b.nx = (int)nx;
  b.ny = (int)ny;
  b.nz = (int)nz;
  b.tx = std::min(b.nx, tfg::kBoxTX);
  b.ty = std::min(b.ny, tfg::kBoxTY);
  b.tz = std::min(b.nz, tfg::kBoxTZ);
  b.ntx = (b.nx + b.tx - 1) / b.tx;
  b.nty = (b.ny + b.ty - 1) / b.ty;
  b.ntz = (b.nz + b.tz - 1) / b.tz;
  std::vector<int> tiles;
  h->box_toff.assign(1, 0);
  for (int l = 0; l <= b.ntx + b.nty + b.ntz - 3; ++l) {
    for (int X = 0; X < b.ntx; ++X)
      for (int Y = 0; Y < b.nty; ++Y) {
        const int Z = l - X - Y;
        if (Z >= 0 && Z < b.ntz) tiles.push_back((X * b.nty + Y) * b.ntz + Z);
      }
    h->box_toff.push_back((int)tiles.size());
  }
  int* dt = nullptr;
  HIP_TRY(hipMalloc(&dt, sizeof(int) * tiles.size()));
  b.tiles = dt;
  HIP_TRY(hipMemcpy(dt, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice));
  h->box = true;
  return TFG_OK;
}

// FP32 sweeps of the symmetric storage: zero-padded slots (the row sums
// and the upper diagonals, refilled per solve) and their row blocks, four
// rows per thread when the offsets allow.
int ensure_sym32(tfg_matrix* h) {
  if (h->sd32) return TFG_OK;
  const int D = h->sdg.off[h->nsym - 1];
  h->stride32 = ((long long)h->ni + D + 4 + 3) & ~3LL;
  bool quad = h->sdg.off[0] == 1 && (h->ni % 4) == 0 && h->nsym <= tfg::kSymPairMax;
  for (int k = 1; k < h->nsym; ++k) quad = quad && (h->sdg.off[k] % 4) == 0;
  std::vector<int> ord = sym_blocks(h->ni, D, quad ? 4 * tfg::kRowsPerBlock : tfg::kRowsPerBlock,
                                    quad ? 4 : 1);
  if (ord.empty() || 4 * h->stride32 * (h->nsym + 1) >= (1LL << 31))
    return fail(TFG_ENOTSUP, "FP32 symmetric storage does not fit");
  const size_t bytes = sizeof(float) * (size_t)(h->nsym + 1) * (size_t)h->stride32;
  HIP_TRY(hipMalloc(&h->sd32, bytes));
  HIP_TRY(hipMemsetAsync(h->sd32, 0, bytes, h->stream));
  HIP_TRY(hipMalloc(&h->sorder32, sizeof(int) * ord.size()));
  HIP_TRY(hipMemcpy(h->sorder32, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
  h->nsorder32 = (int)ord.size() / 2;
  h->quad32 = quad ? 1 : 0;
  const int per_cu = std::max(1, tfg::spmv_blocks_per_cu(tfg::K_SYM_DIA));
  int g = std::min(h->nsorder32, std::max(1, h->cu_count) * per_cu);
  if (g >= tfg::kXcds) g -= g % tfg::kXcds;
  h->grid32 = std::max(1, std::min(g, h->pstride));
  return TFG_OK;
}

}  // namespace

namespace tfg {

// FP32 sweeps inside FP64 refinement (kernels in tfg_ir.hip): the CG of
// Cg.f90 in FP32 on the row-excess matrix, with the constant vector
// deflated, and FP64 refinements (true residual, stopping test, restart)
// decided on the device every chunk of `check` iterations.  niter = FP32
// iterations; fin_res = the criterion on the last FP64 true residual.
int solve_ir(tfg_matrix* h, int prec, const double* val, const double* b, int miter,
             int* niter, double tol, double* fin_res, const double* norm,
             const tfg_options& o, tfg_report* rep) {
  const size_t n = (size_t)h->nt;
  { const int rt = ensure_tuned(h, val); if (rt != TFG_OK) return rt; }
  for (float** v : {&h->ad32, &h->s32, &h->r32, &h->e32, &h->p32, &h->q32})
    if (!*v) HIP_TRY(hipMalloc(v, sizeof(float) * std::max<size_t>(n, 4)));
  if (!h->s64) HIP_TRY(hipMalloc(&h->s64, sizeof(double) * std::max<size_t>(n, 4)));
  // the FP32 product: the symmetric storage when the matrix offers it and
  // val is symmetric bit for bit, else a CSR engine on val32
  bool sym = false;
  if (h->nsym > 0) {
    const int rs = sym_prepare(h, val, &sym);
    if (rs != TFG_OK) return rs;
  }
  if (sym) {
    const int r3 = ensure_sym32(h);
    if (r3 != TFG_OK) return r3;
  } else if (!h->val32) {
    HIP_TRY(hipMalloc(&h->val32, sizeof(float) * (size_t)std::max(h->nnz, 4)));
  }
  tfg::Launch L32 = launch_of(h, h->val32, h->kind);
  if (sym) {
    L32.kind = tfg::K_SYM_DIA;
    L32.grid_spmv = h->grid32;
  } else {
    if (L32.kind != tfg::K_STAGE2048 && L32.kind != tfg::K_PIPE2048_CODED &&
        L32.kind != tfg::K_PIPE_PATTERN)
      L32.kind = tfg::K_PIPE2048;
    L32.grid_spmv = h->grid[L32.kind];
    L32.vec_val = h->nnz >= 4;
  }
  // one grid for every product-shaped kernel of this solve (their partials
  // are swept with it)
  tfg::Launch L = launch_of(h, val, h->kind);
  L.grid_spmv = L32.grid_spmv;
  tfg::CgArgs a = args_of(h, h->kind);
  a.val = val;
  a.b = b;
  a.x = h->x;
  a.r = h->r;
  a.p = h->p;
  a.q = h->q;
  a.ad = h->ad;
  a.pt = partials_of(h, L);
  a.st = h->st;
  tfg::IrArgs f{};
  f.n = h->ni;
  f.val32 = h->val32;
  f.sd32 = h->sd32;
  f.stride32 = h->stride32;
  f.sorder32 = h->sorder32;
  f.nsorder32 = h->nsorder32;
  f.quad = h->quad32;
  f.s64 = h->s64;
  f.s32 = h->s32;
  f.rd32 = h->ad32;
  f.r32 = h->r32;
  f.p32 = h->p32;
  f.q32 = h->q32;
  f.e32 = h->e32;
  f.prec = prec == tfg::PREC_JACOBI ? tfg::PREC_JACOBI : tfg::PREC_NONE;
  if (o.deflation < TFG_DEFLATE_AUTO || o.deflation > TFG_DEFLATE_ON)
    return fail(TFG_EINVAL, "unknown deflation mode %d", o.deflation);
  // deflation needs symmetric values: the symmetric storage checked them;
  // a CSR matrix is checked against its transpose map
  bool check_sym = false;
  if (!sym && o.deflation != TFG_DEFLATE_OFF) {
    const int rt = ensure_tpos(h);
    if (rt != TFG_OK) return rt;
    check_sym = h->tpos_state > 0;
  }

  State* s0 = &h->hst[0];
  memset(s0, 0, sizeof(State));
  s0->tol = tol;
  s0->criterion = o.criterion;
  s0->has_norm = norm ? 1 : 0;
  s0->norm = norm ? *norm : 0.0;
  s0->kcur = 1;
  s0->refine = 1;   // the start runs the restart kernels
  s0->miter = std::max(miter, 0);
  s0->delta = o.inner_tol > 0.0 ? o.inner_tol : 1e-3;
  s0->force = std::max(o.inner_max, 0);
  s0->defl_mode = o.deflation;
  s0->sym_vals = (sym || check_sym) ? 1 : 0;
  HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof(State), hipMemcpyHostToDevice, h->stream));

  const bool timing = o.timing != 0;
  const int max_samples = 256;
  if (timing && h->ev_sample.empty()) {
    h->ev_sample.resize(2 * max_samples, nullptr);
    for (auto& e : h->ev_sample) HIP_TRY(hipEventCreate(&e));
  }
  std::vector<int> sample_k;
  const int ipr = f.prec;
  const bool nrms = o.criterion != TFG_CRIT_RELRES;
  // the deflated update reads s; when deflation is off for certain, it
  // need not
  const bool defl_kernels = o.deflation != TFG_DEFLATE_OFF && (sym || check_sym);
  // ICC: z = M^-1 r from the FP64 DIC sweeps of the widened FP32 residual
  // (the reference's factor, formed once per solve), rounded to FP32
  const bool icc = prec == tfg::PREC_ICC;
  tfg::IccArgs ic{};
  tfg::IrArgs fz = f;
  if (icc) {
    const int ri = ensure_icc(h);
    if (ri != TFG_OK) return ri;
    ic = icc_args(h);
    if (!h->icc_r64) HIP_TRY(hipMalloc(&h->icc_r64, sizeof(double) * std::max<size_t>(n, 4)));
    if (!h->z32) HIP_TRY(hipMalloc(&h->z32, sizeof(float) * std::max<size_t>(n, 4)));
    fz.r32 = h->z32;  // the p update reads z where it reads r otherwise
  }

  HIP_TRY(hipEventRecord(h->ev_t0, h->stream));
  HIP_TRY(tfg::launch_diag_sum(L, a));
  HIP_TRY(tfg::launch_prologue(L, a, ipr));
  HIP_TRY(tfg::launch_start(L, a));
  if (icc) HIP_TRY(tfg::launch_icc_form(h->stream, ic, val, h->st));
  HIP_TRY(tfg::launch_ir_convert(L, L32, a, f, h->nnz));
  if (check_sym)
    HIP_TRY(tfg::launch_ir_symcheck(h->stream, L.grid_vec, val, h->tpos, h->nnz, h->st));
  HIP_TRY(tfg::launch_ir_restart(L, a, f));

  const int check = 8;  // refinement decisions at chunk ends
  int k = 1, slot = 0, prev = -1;
  bool stop = false;
  while (k <= miter && !stop) {
    const int kend = std::min(miter, k + check - 1);
    for (; k <= kend; ++k) {
      if (icc) {
        HIP_TRY(tfg::launch_ir_widen(L, a, f.r32, h->icc_r64));
        HIP_TRY(tfg::launch_icc_apply(h->stream, ic, val, h->st, h->icc_r64, h->icc_z));
        HIP_TRY(tfg::launch_ir_icc_rho(L, a, f, h->icc_z, h->z32));
        HIP_TRY(tfg::launch_ir_pupdate(L, a, fz));
      } else {
        HIP_TRY(tfg::launch_ir_pupdate(L, a, f));
      }
      const bool take = timing && (k % 4 == 1) && (int)sample_k.size() < max_samples;
      if (take) HIP_TRY(hipEventRecord(h->ev_sample[2 * sample_k.size()], h->stream));
      HIP_TRY(tfg::launch_ir_spmv(L32, a, f));
      if (take) {
        HIP_TRY(hipEventRecord(h->ev_sample[2 * sample_k.size() + 1], h->stream));
        sample_k.push_back(k);
      }
      HIP_TRY(tfg::launch_ir_update(L, a, f, defl_kernels, nrms));
    }
    HIP_TRY(tfg::launch_ir_refine(L, a, f));
    HIP_TRY(hipMemcpyAsync(&h->hst[1 + slot], h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipEventRecord(h->ev_poll[slot], h->stream));
    if (prev >= 0) {
      HIP_TRY(hipEventSynchronize(h->ev_poll[prev]));
      if (h->hst[1 + prev].done) stop = true;
    }
    prev = slot;
    slot ^= 1;
  }
  State* sf = &h->hst[3];
  HIP_TRY(hipMemcpyAsync(sf, h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  int status = sf->status;
  if (status == TFG_EBREAKDOWN) {  // the correction since the last refinement
    HIP_TRY(tfg::launch_ir_fold(L, a, f));
  }
  HIP_TRY(hipEventRecord(h->ev_t1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev_t1));
  *niter = sf->done ? sf->niter : std::max(miter, 0);
  *fin_res = sf->res;
  if (rep) {
    rep->status = status;
    rep->iterations = *niter;
    rep->refinements = sf->nrefine;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
    rep->solve_ms = ms;
    double tot = 0.0;
    int cnt = 0;
    for (size_t q = 0; q < sample_k.size(); ++q) {
      if (sample_k[q] > *niter) break;
      float e = 0.f;
      HIP_TRY(hipEventElapsedTime(&e, h->ev_sample[2 * q], h->ev_sample[2 * q + 1]));
      tot += e;
      ++cnt;
    }
    rep->spmv_ms = cnt ? tot / cnt : 0.0;
    rep->spmv_samples = cnt;
    rep->deflated = sf->defl_e > 0.0 ? 1 : 0;
  }
  if (o.true_residual) {
    HIP_TRY(tfg::launch_true_residual(L, a));
    HIP_TRY(hipMemcpyAsync(sf, h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (rep) rep->true_relres = sf->aux0;
  }
  return status;
}

// FP32 sweeps of the textbook methods inside FP64 refinement (kernels in
// tfg_ir_ns.hip): each refinement restarts the method in FP32 on the FP64
// true residual and runs it until its residual has fallen by inner_tol
// (default 1e-4; every inner_max iterations at the latest when > 0), then
// x += e in FP64, the true residual and the stopping test.  The host waits
// at every chunk of 8 inner iterations (the inner solve's length is decided
// there).  niter = FP32 iterations; fin_res = the criterion on the last
// FP64 true residual.
int solve_ir_ns(tfg_matrix* h, int method, int prec, const double* val, const double* b,
                int miter, int* niter, double tol, double* fin_res, const double* norm,
                const tfg_options& o, tfg_report* rep) {
  const size_t n = (size_t)h->nt;
  { const int rt = ensure_tuned(h, val); if (rt != TFG_OK) return rt; }
  for (float** v : {&h->ad32, &h->s32, &h->r32, &h->e32, &h->p32, &h->q32, &h->ns32[0],
                    &h->ns32[1], &h->ns32[2], &h->ns32[3], &h->ns32[4]})
    if (!*v) HIP_TRY(hipMalloc(v, sizeof(float) * std::max<size_t>(n, 4)));
  if (!h->s64) HIP_TRY(hipMalloc(&h->s64, sizeof(double) * std::max<size_t>(n, 4)));
  if (!h->val32) HIP_TRY(hipMalloc(&h->val32, sizeof(float) * (size_t)std::max(h->nnz, 4)));
  if (method == tfg::NS_BICG) {
    { const int rt = ensure_tpos(h); if (rt != TFG_OK) return rt; }
    if (h->tpos_state < 0)
      return fail(TFG_ENOTSUP, "textbook BiCG needs a structurally symmetric matrix");
    if (!h->valt32) HIP_TRY(hipMalloc(&h->valt32, sizeof(float) * (size_t)std::max(h->nnz, 4)));
  }
  tfg::Launch L32 = launch_of(h, h->val32, h->kind);
  if (L32.kind != tfg::K_STAGE2048 && L32.kind != tfg::K_PIPE2048_CODED &&
      L32.kind != tfg::K_PIPE_PATTERN)
    L32.kind = tfg::K_PIPE2048;
  L32.grid_spmv = h->grid[L32.kind];
  L32.vec_val = h->nnz >= 4;
  tfg::Launch L = launch_of(h, val, h->kind);
  L.grid_spmv = L32.grid_spmv;
  tfg::CgArgs a = args_of(h, h->kind);
  a.val = val;
  a.b = b;
  a.x = h->x;
  a.r = h->r;
  a.p = h->p;
  a.q = h->q;
  a.ad = h->ad;
  a.pt = partials_of(h, L);
  a.st = h->st;
  tfg::IrArgs f{};
  f.n = h->ni;
  f.val32 = h->val32;
  f.s64 = h->s64;
  f.s32 = h->s32;
  f.rd32 = h->ad32;
  f.r32 = h->r32;
  f.p32 = h->p32;
  f.q32 = h->q32;
  f.e32 = h->e32;
  f.prec = prec == tfg::PREC_JACOBI ? tfg::PREC_JACOBI : tfg::PREC_NONE;
  tfg::NsArgs32 s{};
  s.valt = h->valt32;
  s.rt = h->ns32[0];
  s.v1 = h->ns32[1];
  s.v2 = h->ns32[2];
  s.v3 = h->ns32[3];
  s.v4 = h->ns32[4];
  tfg::CgArgs ai = a;  // the inner solve's partial grids
  ai.pt.g_upd = method == tfg::NS_CGS ? L32.grid_spmv : L.grid_vec;
  const double tol_in = o.inner_tol > 0.0 ? o.inner_tol : 1e-4;

  State* s0 = &h->hst[0];
  memset(s0, 0, sizeof(State));
  s0->tol = tol;
  s0->criterion = o.criterion;
  s0->has_norm = norm ? 1 : 0;
  s0->norm = norm ? *norm : 0.0;
  s0->refine = 1;   // the FP64 residual kernel runs at every refinement
  HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof(State), hipMemcpyHostToDevice, h->stream));
  State* sf = &h->hst[3];
  State* si = &h->hst[4];
  const bool timing = o.timing != 0;
  const int max_samples = 256;
  if (timing && h->ev_sample.empty()) {
    h->ev_sample.resize(2 * max_samples, nullptr);
    for (auto& e : h->ev_sample) HIP_TRY(hipEventCreate(&e));
  }
  std::vector<int> sample_k;
  auto read = [&](State* dst, const State* src) -> int {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(State), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return TFG_OK;
  };
#define TFG_NS(expr) do { const int rc_ = (expr); if (rc_ != TFG_OK) return rc_; } while (0)

  HIP_TRY(hipEventRecord(h->ev_t0, h->stream));
  HIP_TRY(tfg::launch_diag_sum(L, a));
  HIP_TRY(tfg::launch_prologue(L, a, f.prec));
  HIP_TRY(tfg::launch_start(L, a));
  HIP_TRY(tfg::launch_ir_convert(L, L32, a, f, h->nnz));
  if (method == tfg::NS_BICG) HIP_TRY(tfg::launch_ns32_valt(L, a, h->tpos, h->valt32));
  TFG_NS(read(sf, h->st));
  int total = 0, status = TFG_OK;
  const int inner_cap = o.inner_max > 0 ? o.inner_max : INT32_MAX;
  while (!sf->done) {
    HIP_TRY(tfg::launch_ns32_restart(L, a, f, s, h->st_in, tol_in, method));
    int k = 1;
    const int kmax = std::min(miter - total, inner_cap);
    bool inner_done = false;
    while (k <= kmax && !inner_done) {
      const int kend = std::min(kmax, k + 7);
      for (; k <= kend; ++k) {
        const bool take = timing && (total + k) % 4 == 1 && (int)sample_k.size() < max_samples;
        hipEvent_t e0 = take ? h->ev_sample[2 * sample_k.size()] : nullptr;
        hipEvent_t e1 = take ? h->ev_sample[2 * sample_k.size() + 1] : nullptr;
        HIP_TRY(tfg::launch_ns32_iteration(L, L32, ai, f, s, h->st_in, method, k, e0, e1));
        if (take) sample_k.push_back(total + k);
      }
      TFG_NS(read(si, h->st_in));
      inner_done = si->done != 0;
    }
    const int steps = si->done ? si->niter : (k - 1);
    total += steps;
    HIP_TRY(tfg::launch_ns32_fold_test(L, a, f, total, std::max(miter, 0)));
    TFG_NS(read(sf, h->st));
    if (sf->done) break;
    // an inner breakdown, or an inner solve that could not take a step
    // (its FP32 residual is zero): no refinement can make progress
    if (si->status == TFG_EBREAKDOWN) {
      status = TFG_EBREAKDOWN;
      break;
    }
    if (steps == 0) break;
  }
  if (sf->status == TFG_EBREAKDOWN) status = TFG_EBREAKDOWN;
  HIP_TRY(hipEventRecord(h->ev_t1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev_t1));
  *niter = sf->done ? sf->niter : total;
  *fin_res = sf->res;
  if (rep) {
    rep->status = status;
    rep->iterations = *niter;
    rep->refinements = sf->nrefine;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
    rep->solve_ms = ms;
    double tot = 0.0;
    int cnt = 0;
    for (size_t q = 0; q < sample_k.size(); ++q) {
      if (sample_k[q] > *niter) break;
      float e = 0.f;
      HIP_TRY(hipEventElapsedTime(&e, h->ev_sample[2 * q], h->ev_sample[2 * q + 1]));
      tot += e;
      ++cnt;
    }
    rep->spmv_ms = cnt ? tot / cnt : 0.0;
    rep->spmv_samples = cnt;
  }
  if (o.true_residual) {
    HIP_TRY(tfg::launch_true_residual(L, a));
    TFG_NS(read(sf, h->st));
    if (rep) rep->true_relres = sf->aux0;
  }
#undef TFG_NS
  return status;
}

// One rank of a multi-rank FP64 CG (Cg.f90 with Comm_Mod's calls): every
// global scalar goes through the caller's allreduce, p's buffer cells through
// its exchange before each product, x's at exit.  ICC is the reference's
// per-rank DIC: Prec_Form / Prec_Solve run over this rank's rows 1..ni only
// (the forward recurrences stop at the first column above the row, which a
// buffer cell always is; the backward one skips buffer columns,
// Prec_Solve.f90:89), so M is block diagonal over the ranks and only
// rho = (z, r) crosses them.
int solve_comm(tfg_matrix* h, int prec, const double* val, const double* b, int miter,
               int* niter, double tol, double* fin_res, const double* norm,
               const tfg_options& o, tfg_report* rep) {
  { const int rt = ensure_tuned(h, val); if (rt != TFG_OK) return rt; }
  if (!h->loc) HIP_TRY(hipMalloc(&h->loc, 8 * sizeof(double)));
  const tfg::Launch L = launch_of(h, val, h->kind);
  tfg::CgArgs a = args_of(h, h->kind);
  a.val = val;
  a.b = b;
  a.x = h->x;
  a.r = h->r;
  a.p = h->p;
  a.q = h->q;
  a.ad = h->ad;
  a.pt = partials_of(h, L);
  a.st = h->st;
  const hipStream_t S = h->stream;
  State* s0 = &h->hst[0];
  memset(s0, 0, sizeof(State));
  s0->tol = tol;
  s0->criterion = o.criterion;
  s0->has_norm = norm ? 1 : 0;
  s0->norm = norm ? *norm : 0.0;
  s0->comm = 1;
  HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof(State), hipMemcpyHostToDevice, S));
  double tot[8];
  // this rank's totals of nv partial rows, then the caller's allreduce
  auto reduce = [&](const double* part, int G, int nv, const int* ops) -> int {
    HIP_TRY(tfg::launch_cm_local(S, part, h->pstride, G, nv, ops, h->loc));
    HIP_TRY(hipMemcpyAsync(tot, h->loc, sizeof(double) * nv, hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    int rops[8];
    for (int q = 0; q < nv; ++q) rops[q] = ops[q] == tfg::OP_SUM ? TFG_SUM : (ops[q] == tfg::OP_MIN ? TFG_MIN : TFG_MAX);
    if (h->allreduce(h->comm_ctx, tot, rops, nv) != 0) return fail(TFG_EINVAL, "allreduce hook failed");
    return TFG_OK;
  };
  auto set_field = [&](double* dfield, double v) -> int {
    HIP_TRY(hipMemcpy(dfield, &v, sizeof(double), hipMemcpyHostToDevice));
    return TFG_OK;
  };
  auto exchange = [&](double* d_x) -> int {
    HIP_TRY(hipStreamSynchronize(S));
    if (h->exchange(h->comm_ctx, d_x) != 0) return fail(TFG_EINVAL, "exchange hook failed");
    return TFG_OK;
  };
#define TFG_CM(expr) do { const int rc_ = (expr); if (rc_ != TFG_OK) return rc_; } while (0)
  const bool icc = prec == tfg::PREC_ICC;
  tfg::IccArgs ic{};
  if (icc) {
    const int ri = ensure_icc(h);
    if (ri != TFG_OK) return ri;
    ic = icc_args(h);
  }
  HIP_TRY(hipEventRecord(h->ev_t0, S));
  // fn = 1 / (sum_a / sum_n) over all ranks                     Cg.f90:62-76
  HIP_TRY(tfg::launch_diag_sum(L, a));
  const int sum1[1] = {tfg::OP_SUM};
  TFG_CM(reduce(a.pt.diag, L.grid_vec, 1, sum1));
  {
    double vals[2] = {tot[0], (double)h->ni};
    const int ops2[2] = {TFG_SUM, TFG_SUM};
    if (h->allreduce(h->comm_ctx, vals + 1, ops2 + 1, 1) != 0) return fail(TFG_EINVAL, "allreduce hook failed");
    TFG_CM(set_field(&h->st->fn, 1.0 / (vals[0] / vals[1])));
  }
  // r0 = b - A x and the start sums                            Cg.f90:95-123
  // (ICC: the products and updates run unpreconditioned, z from the sweeps)
  const int pk = icc ? tfg::PREC_NONE : prec;
  HIP_TRY(tfg::launch_prologue(L, a, pk));
  if (icc) HIP_TRY(tfg::launch_icc_form(S, ic, val, h->st));         // Cg.f90:89
  const int pro_ops[7] = {tfg::OP_SUM, tfg::OP_SUM, tfg::OP_MIN, tfg::OP_MAX,
                          tfg::OP_SUM, tfg::OP_SUM, tfg::OP_SUM};
  TFG_CM(reduce(a.pt.pro, L.grid_spmv, 7, pro_ops));
  const double range0 = norm ? std::max(*norm - 0.0, tfg::kTiny) : std::max(tot[3] - tot[2], tfg::kTiny);
  const double bnorm = std::sqrt(tot[4]);
  double bn, res;
  if (o.criterion == TFG_CRIT_RELRES) {
    bn = bnorm > 0.0 ? 1.0 : 0.0;
    res = std::sqrt(tot[5]) / bnorm;
  } else {
    bn = std::sqrt(tot[0]) / range0;
    res = std::sqrt(tot[1]) / range0;
  }
  int iters = 0, status = TFG_OK;
  double rho = tot[6];
  if (bn < tol) {
    res = 0.0;
  } else if (!(res < tol)) {
    const int upd_ops[5] = {tfg::OP_SUM, tfg::OP_MIN, tfg::OP_MAX, tfg::OP_SUM, tfg::OP_SUM};
    int k = 1;
    for (; k <= miter; ++k) {
      if (icc) {
        HIP_TRY(tfg::launch_icc_apply(S, ic, val, h->st, h->r, h->icc_z));  // :145
        HIP_TRY(tfg::launch_icc_rho(L, a, h->icc_z));
        TFG_CM(reduce(a.pt.ext, L.grid_vec, 1, sum1));                 // :150-156
        if (k > 1) TFG_CM(set_field(&h->st->beta, tot[0] / rho));     // :165
        rho = tot[0];
      }
      HIP_TRY(tfg::launch_cm_pupdate(L, a, prec, h->icc_z, k));        // :145-171
      TFG_CM(exchange(h->p));                                          // :176
      HIP_TRY(tfg::launch_spmv_pq(L, a, k));                           // :177-195
      TFG_CM(reduce(a.pt.spmv, L.grid_spmv, 1, sum1));                 // :197
      const double pq = tot[0];
      if (pq == 0.0 || !std::isfinite(pq)) { status = TFG_EBREAKDOWN; break; }
      TFG_CM(set_field(&h->st->alpha, rho / pq));                      // :198
      HIP_TRY(tfg::launch_cm_update(L, a, pk));                        // :204-209
      TFG_CM(reduce(a.pt.upd, L.grid_vec, 5, upd_ops));                // NRMS :38,53-54
      const double range = norm ? std::max(*norm - 0.0, tfg::kTiny) : std::max(tot[2] - tot[1], tfg::kTiny);
      res = o.criterion == TFG_CRIT_RELRES ? std::sqrt(tot[4]) / bnorm : std::sqrt(tot[0]) / range;
      if (!std::isfinite(res)) { status = TFG_EBREAKDOWN; break; }
      if (res < tol) break;                                            // :220
      if (!icc) {
        const double rho_new = tot[3];                                 // :150-156 of k+1
        TFG_CM(set_field(&h->st->beta, rho_new / rho));                // :165
        rho = rho_new;
      }
    }
    iters = std::min(k, miter);
  }
  TFG_CM(exchange(h->x));                                              // :240
  HIP_TRY(hipEventRecord(h->ev_t1, S));
  HIP_TRY(hipEventSynchronize(h->ev_t1));
#undef TFG_CM
  *niter = iters;
  *fin_res = res;
  if (rep) {
    rep->status = status;
    rep->iterations = iters;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
    rep->solve_ms = ms;
  }
  if (o.true_residual) {
    HIP_TRY(tfg::launch_true_residual(L, a));
    State* sf = &h->hst[3];
    HIP_TRY(hipMemcpyAsync(sf, h->st, sizeof(State), hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    if (rep) rep->true_relres = sf->aux0;  // this rank's rows only
  }
  return status;
}

}  // namespace tfg

// ===========================================================================
extern "C" {

// set while tfg_matrix_reorder creates its inner handle (no automatic
// ordering of that one)
static thread_local bool g_inner_create = false;

struct tfg_faces {
  tfg_matrix* h = nullptr;
  int device = 0;            // of h, kept so that destroying needs no h
  int n_faces = 0;
  int n_ambient = 0;         // boundary faces of ambient regions
  int *cf_off = nullptr, *cf_face = nullptr, *ent_face = nullptr, *c1 = nullptr, *c2 = nullptr;
  uint8_t* flags = nullptr;
  int* c2b = nullptr;        // c2 as given (1-based, boundary cells < 0)
  int max_bnd = 0;           // largest boundary cell number
  bool shared_entry = false; // some matrix entry set by two faces
  bool shared_bnd = false;   // some boundary cell on two faces
  double *pmin = nullptr, *pmax = nullptr;  // momentum scratch (nt each), on first use
};

// hipFree waits for the device, so no stream of the matrix handle (which
// may be gone already) is needed; errors of the release are dropped here
// and cleared, so that no later call reports them as its own
static void faces_free(tfg_faces* f) {
  if (!f) return;
  (void)hipSetDevice(f->device);
  void* dev[] = {f->cf_off, f->cf_face, f->ent_face, f->c1, f->c2, f->flags, f->c2b, f->pmin,
                 f->pmax};
  for (void* d : dev)
    if (d) (void)hipFree(d);
  (void)hipGetLastError();
  delete f;
}

int tfg_faces_create(tfg_matrix* h, int n_faces, const int* c1, const int* c2, const int* pos,
                     const int* diag_bc, const int* b_region, tfg_faces** faces) {
  if (!h || !faces || !c1 || !c2 || !pos || !diag_bc || !b_region || n_faces < 0)
    return fail(TFG_EINVAL, "null argument");
  *faces = nullptr;
  const int n = h->nt;
  // host copies of the structure (0-based)
  std::vector<int> row(n + 1), col(h->nnz);
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpy(row.data(), h->row, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(col.data(), h->col, sizeof(int) * (size_t)h->nnz, hipMemcpyDeviceToHost));
  std::vector<int> cnt(n + 1, 0), ent(h->nnz, -1), a(n_faces), b(n_faces);
  std::vector<uint8_t> fl(n_faces);
  bool shared_entry = false;
  int max_bnd = 0;
  for (int s = 0; s < n_faces; ++s)
    if (c2[s] < 0) max_bnd = std::max(max_bnd, -c2[s]);
  std::vector<char> bnd_seen((size_t)max_bnd + 1, 0);
  bool shared_bnd = false;
  int n_ambient = 0;
  for (int s = 0; s < n_faces; ++s) {
    const int p = c1[s] - 1, q = c2[s] > 0 ? c2[s] - 1 : -1;
    if (p < 0 || p >= n || q >= n) return fail(TFG_EINVAL, "face %d: cells out of range", s + 1);
    if (b_region[s] < TFG_REGION_NONE || b_region[s] > TFG_REGION_AMBIENT)
      return fail(TFG_EINVAL, "face %d: region code %d", s + 1, b_region[s]);
    a[s] = p;
    b[s] = q;
    uint8_t code = 0;
    if (q < 0 && c2[s] < 0) {   // a boundary cell: its region's assignment of b(c1)
      code = b_region[s] == TFG_REGION_PRESSURE ? 2 : (b_region[s] == TFG_REGION_AMBIENT ? 4 : 0);
      if (code == 4) ++n_ambient;
    }
    fl[s] = (uint8_t)(q < 0 ? ((diag_bc[s] ? 1 : 0) | code) : 0);
    cnt[p + 1]++;
    if (q >= 0) {
      cnt[q + 1]++;
      const int j1 = pos[2 * s] - 1, j2 = pos[2 * s + 1] - 1;
      if (j1 <= row[p] || j1 >= row[p + 1] || col[j1] != q || j2 <= row[q] || j2 >= row[q + 1] ||
          col[j2] != p)
        return fail(TFG_EINVAL, "face %d: pos does not point at (c1,c2) / (c2,c1)", s + 1);
      if (ent[j1] >= 0 || ent[j2] >= 0) shared_entry = true;
      ent[j1] = s;   // the last face that sets the entry wins, as in the loop
      ent[j2] = s;
    } else if (c2[s] < 0) {
      if (bnd_seen[-c2[s]]) shared_bnd = true;
      bnd_seen[-c2[s]] = 1;
    }
  }
  for (int c = 0; c < n; ++c) cnt[c + 1] += cnt[c];
  std::vector<int> cf(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
  for (int s = 0; s < n_faces; ++s) {   // increasing face order within each cell
    cf[fill[a[s]]++] = s << 1;
    if (b[s] >= 0) cf[fill[b[s]]++] = (s << 1) | 1;
  }
  tfg_faces* f = new tfg_faces();
  f->h = h;
  f->device = h->device;
  f->n_faces = n_faces;
  hipError_t e = hipSuccess;
  auto up = [&](void** d, const void* src, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(d, bytes > 0 ? bytes : 4);
    if (e == hipSuccess && bytes) e = hipMemcpy(*d, src, bytes, hipMemcpyHostToDevice);
  };
  up((void**)&f->cf_off, cnt.data(), sizeof(int) * cnt.size());
  up((void**)&f->cf_face, cf.data(), sizeof(int) * cf.size());
  up((void**)&f->ent_face, ent.data(), sizeof(int) * ent.size());
  up((void**)&f->c1, a.data(), sizeof(int) * a.size());
  up((void**)&f->c2, b.data(), sizeof(int) * b.size());
  up((void**)&f->flags, fl.data(), fl.size());
  up((void**)&f->c2b, c2, sizeof(int) * (size_t)n_faces);
  f->max_bnd = max_bnd;
  f->n_ambient = n_ambient;
  f->shared_entry = shared_entry;
  f->shared_bnd = shared_bnd;
  if (e != hipSuccess) {
    faces_free(f);
    return fail(TFG_ENOMEM, "face structure upload: %s", hipGetErrorString(e));
  }
  *faces = f;
  return TFG_OK;
}

int tfg_momentum_assemble(tfg_faces* f, const tfg_momentum* m, double* val, double* b,
                          double* v_m) {
  if (!f || !m || !val || !b || !v_m) return fail(TFG_EINVAL, "null argument");
  const double* need[] = {m->f, m->fw, m->fc, m->v_flux, m->vis_eff, m->sx, m->sy, m->sz, m->dx,
                          m->dy, m->dz, m->phi_n, m->phi_o, m->phi_oo, m->gx, m->gy, m->gz,
                          m->uj_i, m->uk_i, m->density, m->vol, m->p_i};
  for (const double* p : need)
    if (!p) return fail(TFG_EINVAL, "null momentum array");
  if (!m->wall) return fail(TFG_EINVAL, "null momentum array");
  if (m->component < 1 || m->component > 3) return fail(TFG_EINVAL, "component %d", m->component);
  if (m->adv_scheme < TFG_ADV_CENTRAL || m->adv_scheme > TFG_ADV_AVL_SMART)
    return fail(TFG_ENOTSUP, "advection scheme %d", m->adv_scheme);
  if (m->td_scheme < TFG_TD_LINEAR || m->td_scheme > TFG_TD_NONE)
    return fail(TFG_EINVAL, "time scheme %d", m->td_scheme);
  if (m->n_bnd < f->max_bnd)
    return fail(TFG_EINVAL, "n_bnd %d below the faces' boundary cell %d", m->n_bnd, f->max_bnd);
  if (f->shared_entry || f->shared_bnd)
    return fail(TFG_ENOTSUP, "a matrix entry or a boundary cell belongs to two faces");
  tfg_matrix* h = f->h;
  HIP_TRY(hipSetDevice(h->device));
  if (m->adv_scheme != TFG_ADV_CENTRAL && !f->pmin) {
    HIP_TRY(hipMalloc(&f->pmin, sizeof(double) * (size_t)std::max(1, h->nt)));
    HIP_TRY(hipMalloc(&f->pmax, sizeof(double) * (size_t)std::max(1, h->nt)));
  }
  tfg::MomArgs g{};
  g.n = h->nt;
  g.nb = m->n_bnd;
  g.row = h->row;
  g.cf_off = f->cf_off;
  g.cf_face = f->cf_face;
  g.ent_face = f->ent_face;
  g.c1 = f->c1;
  g.c2b = f->c2b;
  g.comp = m->component;
  g.adv = m->adv_scheme;
  g.blend_matrix = m->blend_matrix ? 1 : 0;
  g.td = m->td_scheme;
  g.blend = m->blend;
  g.dt = m->dt;
  g.urf = m->urf;
  g.p_drop = m->p_drop;
  g.f = m->f;
  g.fw = m->fw;
  g.fc = m->fc;
  g.v_flux = m->v_flux;
  g.vis_eff = m->vis_eff;
  g.f_stress = m->f_stress;
  g.sv[0] = m->sx;
  g.sv[1] = m->sy;
  g.sv[2] = m->sz;
  g.dv[0] = m->dx;
  g.dv[1] = m->dy;
  g.dv[2] = m->dz;
  g.wall = m->wall;
  g.phi = m->phi_n;
  g.phi_o = m->phi_o;
  g.phi_oo = m->phi_oo;
  g.g[0] = m->gx;
  g.g[1] = m->gy;
  g.g[2] = m->gz;
  g.uj_i = m->uj_i;
  g.uk_i = m->uk_i;
  g.rho = m->density;
  g.vol = m->vol;
  g.p_i = m->p_i;
  g.cell_f = m->cell_f;
  g.st = m->st;
  g.pmin = f->pmin;
  g.pmax = f->pmax;
  g.val = val;
  g.b = b;
  g.v_m = v_m;
  HIP_TRY(tfg::launch_momentum(h->stream, g));
  return TFG_OK;
}

int tfg_rhie_chow_fluxes(tfg_faces* f, const tfg_rhie_chow* r, double* v_flux) {
  if (!f || !r || !v_flux) return fail(TFG_EINVAL, "null argument");
  const double* need[] = {r->f, r->fc, r->sx, r->sy, r->sz, r->dx, r->dy, r->dz, r->u, r->v, r->w,
                          r->p, r->px, r->py, r->pz, r->v_m};
  for (const double* q : need)
    if (!q) return fail(TFG_EINVAL, "null Rhie-Chow array");
  if (r->choi) {
    const double* choi[] = {r->u_o, r->u_oo, r->v_o, r->v_oo, r->w_o_, r->w_oo_, r->density,
                            r->flux_o, r->flux_oo};
    for (const double* q : choi)
      if (!q) return fail(TFG_EINVAL, "null array of Choi's correction");
  }
  if (r->gu) {
    const double* gu[] = {r->cell_fx, r->cell_fy, r->cell_fz, r->face_fx, r->face_fy, r->face_fz};
    for (const double* q : gu)
      if (!q) return fail(TFG_EINVAL, "null array of Gu's correction");
  }
  const bool vof = r->vof && r->sigma > 1.0e-15;   // Rhie_And_Chow.f90:104-105, TINY
  if (vof) {
    const double* st[] = {r->curv, r->col, r->col_x, r->col_y, r->col_z};
    for (const double* q : st)
      if (!q) return fail(TFG_EINVAL, "null surface-tension array");
  }
  HIP_TRY(hipSetDevice(f->device));
  tfg::RcArgs a{};
  a.n_faces = f->n_faces;
  a.c1 = f->c1;
  a.c2 = f->c2;
  a.r = *r;
  a.vof = vof ? 1 : 0;
  a.v_flux = v_flux;
  HIP_TRY(tfg::launch_rhie_chow(f->h->stream, a));
  return TFG_OK;
}

int tfg_faces_destroy(tfg_faces* f) {
  if (!f) return fail(TFG_EINVAL, "null handle");
  faces_free(f);
  return TFG_OK;
}

int tfg_pressure_assemble(tfg_faces* f, const double* fc, const double* fw, const double* vm,
                          const double* vflux, double* val, double* b, int singular) {
  return tfg_pressure_assemble_ex(f, fc, fw, vm, vflux, nullptr, val, b, singular);
}

int tfg_pressure_assemble_ex(tfg_faces* f, const double* fc, const double* fw, const double* vm,
                             const double* vflux, const double* pp, double* val, double* b,
                             int singular) {
  if (!f || !fc || !fw || !vm || !vflux || !val || !b) return fail(TFG_EINVAL, "null argument");
  if (f->n_ambient > 0 && !pp)
    return fail(TFG_EINVAL, "%d faces of ambient regions need pp % n", f->n_ambient);
  tfg_matrix* h = f->h;
  HIP_TRY(hipSetDevice(h->device));
  tfg::AsmArgs g{};
  g.n = h->nt;
  g.row = h->row;
  g.cf_off = f->cf_off;
  g.cf_face = f->cf_face;
  g.ent_face = f->ent_face;
  g.c1 = f->c1;
  g.c2 = f->c2;
  g.flags = f->flags;
  g.c2b = f->c2b;
  g.pp = pp;
  g.fc = fc;
  g.fw = fw;
  g.vm = vm;
  g.vflux = vflux;
  g.val = val;
  g.b = b;
  g.singular = singular ? 1 : 0;
  HIP_TRY(tfg::launch_assemble(h->stream, g));
  return TFG_OK;
}

int tfg_vel_max(tfg_matrix* h, const double* u, const double* v, const double* w,
                double* vel_max) {
  if (!h || !u || !v || !w || !vel_max) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  const int G = std::max(1, std::min((h->nt + tfg::kBlock - 1) / tfg::kBlock, 1024));
  if (!h->vmax) HIP_TRY(hipMalloc(&h->vmax, sizeof(double) * 1025));
  HIP_TRY(tfg::launch_vel_max(h->stream, h->nt, u, v, w, h->vmax, G, h->vmax + 1024));
  HIP_TRY(hipMemcpyAsync(vel_max, h->vmax + 1024, sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return TFG_OK;
}

int tfg_range(tfg_matrix* h, const double* v, double* vmin, double* vmax) {
  if (!h || !v || !vmin || !vmax) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  const int G = std::max(1, std::min((h->nt + tfg::kBlock - 1) / tfg::kBlock, 1024));
  if (!h->vrange) HIP_TRY(hipMalloc(&h->vrange, sizeof(double) * (2 * 1024 + 2)));
  HIP_TRY(tfg::launch_range(h->stream, h->nt, v, h->vrange, G, h->vrange + 2 * 1024));
  double out[2];
  HIP_TRY(hipMemcpyAsync(out, h->vrange + 2 * 1024, sizeof(out), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *vmin = out[0];
  *vmax = out[1];
  return TFG_OK;
}

int tfg_matrix_set_comm(tfg_matrix* h, tfg_allreduce_fn allreduce, tfg_exchange_fn exchange,
                        void* ctx) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if ((allreduce == nullptr) != (exchange == nullptr))
    return fail(TFG_EINVAL, "allreduce and exchange hooks go together");
  if (allreduce && h->inner) {   // the multi-rank path runs in the caller's order
    const int rc = tfg_matrix_reorder(h, TFG_ORDER_NATURAL);
    if (rc != TFG_OK) return rc;
  }
  h->allreduce = allreduce;
  h->exchange = exchange;
  h->comm_ctx = ctx;
  return TFG_OK;
}

const char* tfg_last_error(void) { return tfg::g_err.c_str(); }

int tfg_version(void) { return 1; }

int tfg_matrix_create(int ni, int nt, int nnz, const int* row, const int* col, const int* dia,
                      tfg_matrix** handle) {
  if (!handle || !row || !col || !dia) return fail(TFG_EINVAL, "null argument");
  *handle = nullptr;
  if (nt < 1 || ni < 1 || ni > nt) return fail(TFG_EINVAL, "bad sizes ni=%d nt=%d", ni, nt);
  { const int rc = tfg::check_csr(nt, nnz, row, col, dia); if (rc != TFG_OK) return rc; }

  tfg_matrix* h = new tfg_matrix();
  h->ni = ni;
  h->nt = nt;
  h->nnz = nnz;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(TFG_EHIP, "no HIP device");
  }
  int rc = TFG_OK;
  do {
    hipError_t e = hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount,
                                         h->device);
    if (e != hipSuccess) { rc = fail(TFG_EHIP, "device attribute: %s", hipGetErrorString(e)); break; }
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { rc = fail(TFG_EHIP, "stream: %s", hipGetErrorString(e)); break; }
    rc = scratch_setup(h->device);
    if (rc != TFG_OK) break;

    std::vector<int> tmp((size_t)std::max(nnz, nt + 1));
    // Row blocks of the two stage engines (make_blocks reads 1-based starts;
    // only differences matter).
    std::vector<int> blk[2];
    // products and sums run over the solved rows 1..ni only
    make_blocks(row, ni, tfg::kCap, blk[0]);
    make_blocks(row, ni, tfg::kCap2, blk[1]);
    e = hipMalloc(&h->row, sizeof(int) * (size_t)(nt + 1));
    if (e == hipSuccess) e = hipMalloc(&h->col, sizeof(int) * (size_t)nnz);
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
      h->nblk[b] = (int)blk[b].size() - 1;
      e = hipMalloc(&h->blk[b], sizeof(int) * blk[b].size());
    }
    if (e != hipSuccess) { rc = fail(TFG_ENOMEM, "structure allocation: %s", hipGetErrorString(e)); break; }
    for (int i = 0; i <= nt; ++i) tmp[i] = row[i] - 1;
    e = hipMemcpy(h->row, tmp.data(), sizeof(int) * (size_t)(nt + 1), hipMemcpyHostToDevice);
    for (long long j = 0; j < nnz; ++j) tmp[j] = col[j] - 1;
    if (e == hipSuccess) e = hipMemcpy(h->col, tmp.data(), sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice);
    for (int b = 0; b < 2 && e == hipSuccess; ++b)
      e = hipMemcpy(h->blk[b], blk[b].data(), sizeof(int) * blk[b].size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { rc = fail(TFG_EHIP, "structure upload: %s", hipGetErrorString(e)); break; }
    // DIC dependency levels (not for the renumbered handle of an internal
    // ordering: ICC runs on the caller's handle, in the caller's order)
    if (!g_inner_create) {
      std::vector<int> r0(row, row + nt + 1), c0(col, col + nnz);
      for (auto& v : r0) v -= 1;
      for (auto& v : c0) v -= 1;
      for (int f = 0; f < 2 && e == hipSuccess; ++f) {
        std::vector<int> rows;
        icc_levels(r0.data(), c0.data(), nt, f == 0, rows, h->icc_off[f]);
        e = hipMalloc(&h->icc_rows[f], sizeof(int) * (size_t)nt);
        if (e == hipSuccess)
          e = hipMemcpy(h->icc_rows[f], rows.data(), sizeof(int) * (size_t)nt, hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) { rc = fail(TFG_EHIP, "level upload: %s", hipGetErrorString(e)); break; }
      rc = box_setup(h, r0.data(), c0.data());
      if (rc != TFG_OK) break;
    }
    // Row patterns when the rows have at most kPatCount distinct offset
    // sequences of at most kPatMax entries (structured grids)
    {
      std::vector<std::vector<int>> seqs;
      std::vector<uint8_t> pat((size_t)nt + 64, 0);
      bool ok = nnz >= 64;
      std::vector<int> cur;
      int last = -1;
      for (int i = 0; i < ni && ok; ++i) {
        const int len = row[i + 1] - row[i];
        if (len > tfg::kPatMax) { ok = false; break; }
        cur.assign(len, 0);
        for (int j = 0; j < len; ++j) cur[j] = col[row[i] - 1 + j] - 1 - i;
        int id = -1;
        if (last >= 0 && seqs[last] == cur) id = last;
        for (size_t t = 0; t < seqs.size() && id < 0; ++t)
          if (seqs[t] == cur) id = (int)t;
        if (id < 0) {
          if ((int)seqs.size() == tfg::kPatCount) { ok = false; break; }
          id = (int)seqs.size();
          seqs.push_back(cur);
        }
        pat[i] = (uint8_t)id;
        last = id;
      }
      if (ok) {
        std::vector<int> poff(tfg::kPatCount * tfg::kPatMax, 0), plen(tfg::kPatCount, 1);
        for (size_t t = 0; t < seqs.size(); ++t) {
          plen[t] = (int)seqs[t].size();
          for (size_t j = 0; j < seqs[t].size(); ++j) poff[t * tfg::kPatMax + j] = seqs[t][j];
        }
        e = hipMalloc(&h->pat, pat.size());
        if (e == hipSuccess) e = hipMalloc(&h->poff, sizeof(int) * poff.size());
        if (e == hipSuccess) e = hipMalloc(&h->plen, sizeof(int) * plen.size());
        if (e == hipSuccess) e = hipMemcpy(h->pat, pat.data(), pat.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->poff, poff.data(), sizeof(int) * poff.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->plen, plen.data(), sizeof(int) * plen.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) { rc = fail(TFG_EHIP, "pattern upload: %s", hipGetErrorString(e)); break; }
        h->npat = (int)seqs.size();
        // Symmetric engine: no buffer cells, the diagonal first in every
        // row and the other columns ascending (the order the engine sums
        // in), at most kSymMax distinct positive offsets, and every entry
        // (i, i + d) has its mirror (i + d, i).
        std::vector<int> up;
        bool sym_ok = ni == nt;
        for (const auto& q : seqs) {
          if (q.empty() || q[0] != 0) sym_ok = false;  // diagonal first
          for (size_t u = 2; u < q.size() && sym_ok; ++u)
            if (!(q[u - 1] < q[u])) sym_ok = false;    // then ascending
          for (size_t u = 0; u < q.size() && sym_ok; ++u) {
            if (std::count(q.begin(), q.end(), q[u]) != 1) sym_ok = false;
            if (q[u] > 0 && std::find(up.begin(), up.end(), q[u]) == up.end()) up.push_back(q[u]);
          }
        }
        if (up.empty() || (int)up.size() > tfg::kSymMax) sym_ok = false;
        std::sort(up.begin(), up.end());
        if (sym_ok) {
          // has[t][k]: pattern t holds offset -up[k]
          std::vector<char> has(seqs.size() * up.size(), 0);
          for (size_t t = 0; t < seqs.size(); ++t)
            for (size_t k = 0; k < up.size(); ++k)
              has[t * up.size() + k] =
                  std::find(seqs[t].begin(), seqs[t].end(), -up[k]) != seqs[t].end();
          std::vector<int> upos(seqs.size() * tfg::kPatMax, -1);  // slot of a positive entry
          for (size_t t = 0; t < seqs.size(); ++t)
            for (size_t u = 0; u < seqs[t].size(); ++u)
              if (seqs[t][u] > 0)
                upos[t * tfg::kPatMax + u] =
                    (int)(std::find(up.begin(), up.end(), seqs[t][u]) - up.begin());
          for (int i = 0; i < ni && sym_ok; ++i) {
            const int t = pat[i];
            for (size_t u = 0; u < seqs[t].size(); ++u) {
              const int k = upos[t * tfg::kPatMax + u];
              if (k < 0) continue;
              const long long j = (long long)i + seqs[t][u];
              if (j >= ni || !has[(size_t)pat[j] * up.size() + k]) { sym_ok = false; break; }
            }
          }
          // lower entries: every (i, i - d) is the mirror of an upper entry
          // of row i - d (same count of entries on both sides)
          long long nup = 0, nlow = 0;
          for (int i = 0; i < ni && sym_ok; ++i)
            for (int d : seqs[pat[i]]) {
              if (d > 0) ++nup;
              else if (d < 0) ++nlow;
            }
          if (nup != nlow) sym_ok = false;
        }
        if (sym_ok) {
          std::vector<int> slot(tfg::kPatCount * tfg::kPatMax, 0);
          for (size_t t = 0; t < seqs.size(); ++t)
            for (size_t u = 0; u < seqs[t].size(); ++u) {
              const int d = seqs[t][u];
              const int k = (int)(std::find(up.begin(), up.end(), d > 0 ? d : -d) - up.begin());
              slot[t * tfg::kPatMax + u] = d == 0 ? 0 : d > 0 ? k + 1 : -(k + 1);
            }
          e = hipMalloc(&h->pslot, sizeof(int) * slot.size());
          if (e == hipSuccess)
            e = hipMemcpy(h->pslot, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice);
          if (e != hipSuccess) { rc = fail(TFG_EHIP, "slot upload: %s", hipGetErrorString(e)); break; }
          h->nsym = (int)up.size();
          // Slot stride: n and then at least D + 2 zeros, even, so that the
          // lower entries of the first rows read zeros.  Two rows per thread
          // when the offsets are 1 and even ones and n is even.
          const int D = up.back();
          const long long gap = ((long long)D + 2 + 1) & ~1LL;
          h->sdg = tfg::SymDiag{};
          h->sdg.stride = ((long long)ni + gap + 1) & ~1LL;
          h->sdg.k = h->nsym;
          for (int k = 0; k < h->nsym; ++k) h->sdg.off[k] = up[k];
          bool pair = up[0] == 1 && (ni % 2) == 0 && h->nsym <= tfg::kSymPairMax;
          for (int k = 1; k < h->nsym; ++k) pair = pair && (up[k] % 2) == 0;
          h->sdg.pair = pair ? 1 : 0;
          // Processing order of the symmetric engine: planes of D rows (D =
          // the largest offset; plane starts rounded down to even rows),
          // blocks of at most R rows (256, or 512 in pair mode) inside a
          // plane, taken band by band (8 consecutive blocks of a plane) and,
          // within a band, plane after plane.  The workgroups that run at
          // the same time then cover one band over neighbouring planes, so
          // most rows a row reads (one plane or one line away) are read by
          // them too and stay in L2: 377 instead of 522 MB fetched per
          // product at 200^3.
          const int R = pair ? 2 * tfg::kRowsPerBlock : tfg::kRowsPerBlock;
          std::vector<int> ord = sym_blocks(ni, D, R, pair ? 2 : 1);
          // (not expected) no symmetric engine; the kernels address a slot
          // with 32-bit byte offsets
          if (ord.empty() || h->sdg.stride >= (1LL << 29)) h->nsym = 0;
          if (h->nsym > 0) {
            e = hipMalloc(&h->sorder, sizeof(int) * ord.size());
            if (e == hipSuccess)
              e = hipMemcpy(h->sorder, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) { rc = fail(TFG_EHIP, "order upload: %s", hipGetErrorString(e)); break; }
            h->nsorder = (int)ord.size() / 2;
          }
        }
      }
    }
    // Offset codes when the matrix has at most 256 distinct col - row
    // offsets (every structured grid): 1 byte per nonzero instead of 4.
    {
      std::vector<int> table;
      std::vector<uint8_t> code((size_t)nnz + 64, 0);
      bool ok = nnz >= 64;
      int last_off = INT_MIN, last_code = 0;
      for (int i = 0; i < nt && ok; ++i) {
        for (int j = row[i] - 1; j < row[i + 1] - 1; ++j) {
          const int off = col[j] - 1 - i;
          if (off != last_off) {
            int c = -1;
            for (size_t t = 0; t < table.size(); ++t)
              if (table[t] == off) { c = (int)t; break; }
            if (c < 0) {
              if (table.size() == 256) { ok = false; break; }
              c = (int)table.size();
              table.push_back(off);
            }
            last_off = off;
            last_code = c;
          }
          code[j] = (uint8_t)last_code;
        }
      }
      if (ok) {
        e = hipMalloc(&h->code, code.size());
        if (e == hipSuccess) e = hipMalloc(&h->tab, sizeof(int) * 256);
        if (e == hipSuccess) e = hipMemcpy(h->code, code.data(), code.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->tab, table.data(), sizeof(int) * table.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) { rc = fail(TFG_EHIP, "code upload: %s", hipGetErrorString(e)); break; }
        h->ntab = (int)table.size();
      }
    }
    rc = alloc_work(h);
    const char* force = getenv("TFG_SPMV_KIND");
    if (rc == TFG_OK && force && *force) {
      const int k = atoi(force);
      if (kind_available(h, k)) h->kind = k;  // else the default choice
    }
    // A large matrix without row patterns whose numbering has no locality
    // (mean |col - row| of the off-diagonals above nt/64; a random numbering
    // gives about nt/3) gets the internal RCM ordering.
    if (rc == TFG_OK && !g_inner_create && h->npat == 0 && ni == nt && nt >= (1 << 16)) {
      double dist = 0.0;
      long long cnt = 0;
      for (int i = 0; i < ni; ++i)
        for (int j = row[i] - 1; j < row[i + 1] - 1; ++j)
          if (col[j] - 1 != i) {
            dist += std::abs(col[j] - 1 - i);
            ++cnt;
          }
      if (cnt > 0 && dist / (double)cnt > nt / 64.0) (void)tfg_matrix_reorder(h, TFG_ORDER_RCM);
    }
  } while (0);
  if (rc != TFG_OK) {
    free_all(h);
    return rc;
  }
  *handle = h;
  return TFG_OK;
}

extern "C++" {  // C++ helpers inside the C-ABI block
namespace {

// Reverse Cuthill-McKee order of the graph of rows (0-based CSR): for each
// connected part, a breadth-first sweep from a pseudo-peripheral row
// (George and Liu: repeat the sweep from a least-degree row of the last
// level while the depth grows), neighbours in increasing degree; the whole
// order reversed.  Returns new -> old.
//
// The search sweeps only need each row's level, so they run level by level
// on up to 16 host threads: a row joins the next level through a compare-
// and-swap of its mark, and which thread wins changes neither the levels
// nor the pick (least degree, then smallest index), so the order is the same
// on every run and for any thread count.  The ordering sweep itself follows
// the queue and stays serial.
std::vector<int> rcm_order(int n, const std::vector<int>& row, const std::vector<int>& col) {
  std::vector<int> deg(n), order, mark(n, 0);  // mark: -1 ordered, else the last search
  for (int i = 0; i < n; ++i) deg[i] = row[i + 1] - row[i];
  std::vector<int> by_deg(n);
  for (int i = 0; i < n; ++i) by_deg[i] = i;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  order.reserve(n);
  const int T = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::vector<int>> next(T);
  std::vector<int> front, nb;
  int search = 0;
  // levels from s over rows not yet ordered; returns the depth and the
  // least-degree row of the last level (smallest index on ties)
  auto levels = [&](int s, int& last) -> int {
    const int id = ++search;
    front.assign(1, s);
    mark[s] = id;
    int depth = 0;
    for (;;) {
      auto part = [&](int t) {
        std::vector<int>& out = next[t];
        out.clear();
        const size_t lo = front.size() * t / T, hi = front.size() * (t + 1) / T;
        for (size_t h = lo; h < hi; ++h) {
          const int u = front[h];
          for (int k = row[u]; k < row[u + 1]; ++k) {
            const int v = col[k];
            int m = __atomic_load_n(&mark[v], __ATOMIC_RELAXED);
            if (m == id || m < 0) continue;
            if (__atomic_compare_exchange_n(&mark[v], &m, id, false, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED))
              out.push_back(v);
          }
        }
      };
      if (T == 1 || front.size() < 4096) {
        for (int t = 0; t < T; ++t) part(t);
      } else {
        // parts whose thread cannot be started run on this one
        std::vector<std::thread> th;
        int started = 1;
        try {
          for (int t = 1; t < T; ++t) {
            th.emplace_back(part, t);
            started = t + 1;
          }
        } catch (...) {
        }
        for (int t = started; t < T; ++t) part(t);
        part(0);
        for (std::thread& x : th) x.join();
      }
      size_t total = 0;
      for (int t = 0; t < T; ++t) total += next[t].size();
      if (total == 0) break;
      ++depth;
      front.clear();
      for (int t = 0; t < T; ++t) front.insert(front.end(), next[t].begin(), next[t].end());
    }
    last = -1;
    for (int u : front)
      if (last < 0 || deg[u] < deg[last] || (deg[u] == deg[last] && u < last)) last = u;
    return depth;
  };
  size_t cand = 0;
  while ((int)order.size() < n) {
    while (mark[by_deg[cand]] < 0) ++cand;
    int s = by_deg[cand], last = -1;
    int depth = levels(s, last);
    for (int it = 0; it < 4; ++it) {
      int last2 = -1;
      const int d2 = levels(last, last2);
      if (d2 <= depth) break;
      s = last;
      depth = d2;
      last = last2;
    }
    const size_t start = order.size();
    order.push_back(s);
    mark[s] = -1;
    for (size_t h = start; h < order.size(); ++h) {
      const int u = order[h];
      nb.clear();
      for (int k = row[u]; k < row[u + 1]; ++k)
        if (mark[col[k]] >= 0) {
          mark[col[k]] = -1;
          nb.push_back(col[k]);
        }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

bool is_icc(const char* prec, int prec_len) { return fstring(prec, prec_len) == "icc"; }

// permuted copies of a value array (by vmap) and vectors (by iperm) on the
// inner handle's stream; back by perm
hipError_t to_inner_val(tfg_matrix* h, const double* val) {
  return tfg::launch_gather(h->inner->stream, val, h->vmap, h->pval, h->nnz);
}
hipError_t to_inner_vec(tfg_matrix* h, const double* v, double* dst) {
  return tfg::launch_gather(h->inner->stream, v, h->iperm, dst, h->nt);
}
hipError_t from_inner_vec(tfg_matrix* h, const double* src, double* v) {
  return tfg::launch_gather(h->inner->stream, src, h->perm, v, h->nt);
}

}  // namespace
}  // extern "C++"

int tfg_matrix_reorder(tfg_matrix* h, int order) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (order != TFG_ORDER_NATURAL && order != TFG_ORDER_RCM)
    return fail(TFG_EINVAL, "unknown ordering %d", order);
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner) {
    free_all(h->inner);
    h->inner = nullptr;
    for (void** d : {(void**)&h->vmap, (void**)&h->perm, (void**)&h->iperm, (void**)&h->pval,
                     (void**)&h->pb, (void**)&h->px, (void**)&h->pq}) {
      if (*d) HIP_TRY(hipFree(*d));
      *d = nullptr;
    }
  }
  if (order == TFG_ORDER_NATURAL) return TFG_OK;
  if (h->ni < h->nt || h->allreduce)
    return fail(TFG_ENOTSUP, "internal ordering with buffer cells or communication hooks");
  const int n = h->nt, nnz = h->nnz;
  std::vector<int> row(n + 1), col(nnz);
  HIP_TRY(hipMemcpy(row.data(), h->row, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(col.data(), h->col, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToHost));
  const std::vector<int> iperm = rcm_order(n, row, col);
  std::vector<int> perm(n);
  for (int k = 0; k < n; ++k) perm[iperm[k]] = k;
  // the renumbered CSR, 1-based: row k is old row iperm[k], entries in the
  // caller's order (diagonal first), columns renumbered
  std::vector<int> row1(n + 1), col1(nnz), dia1(n), vmap(nnz);
  row1[0] = 1;
  for (int k = 0; k < n; ++k) {
    const int o = iperm[k];
    int w = row1[k] - 1;
    for (int j = row[o]; j < row[o + 1]; ++j, ++w) {
      col1[w] = perm[col[j]] + 1;
      vmap[w] = j;
    }
    row1[k + 1] = w + 1;
    dia1[k] = row1[k];
  }
  tfg_matrix* in = nullptr;
  g_inner_create = true;
  const int rc = tfg_matrix_create(n, n, nnz, row1.data(), col1.data(), dia1.data(), &in);
  g_inner_create = false;
  if (rc != TFG_OK) return rc;
  h->inner = in;
  auto up = [&](int** d, const std::vector<int>& v) -> hipError_t {
    hipError_t e = hipMalloc(d, sizeof(int) * std::max<size_t>(1, v.size()));
    if (e == hipSuccess) e = hipMemcpy(*d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = up(&h->vmap, vmap);
  if (e == hipSuccess) e = up(&h->perm, perm);
  if (e == hipSuccess) e = up(&h->iperm, iperm);
  if (e == hipSuccess) e = hipMalloc(&h->pval, sizeof(double) * (size_t)std::max(1, nnz));
  for (double** d : {&h->pb, &h->px, &h->pq})
    if (e == hipSuccess) e = hipMalloc(d, sizeof(double) * (size_t)std::max(1, n));
  if (e != hipSuccess) {
    (void)tfg_matrix_reorder(h, TFG_ORDER_NATURAL);
    return fail(TFG_ENOMEM, "ordering buffers: %s", hipGetErrorString(e));
  }
  return TFG_OK;
}

int tfg_matrix_destroy(tfg_matrix* h) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  free_all(h);
  return TFG_OK;
}

int tfg_matrix_info(tfg_matrix* h, int* n, int* nnz, int* spmv_kind, int* spmv_grid,
                    int* stage_blocks, int* offsets, int* patterns, int* sym_diagonals,
                    int* sym_state, int* ordering, int* icc_tiles) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (ordering) *ordering = h->inner ? TFG_ORDER_RCM : TFG_ORDER_NATURAL;
  // ICC runs in the caller's order, on this handle
  if (icc_tiles) *icc_tiles = h->box ? h->bx.ntx * h->bx.nty * h->bx.ntz : 0;
  if (h->inner)
    return tfg_matrix_info(h->inner, n, nnz, spmv_kind, spmv_grid, stage_blocks, offsets,
                           patterns, sym_diagonals, sym_state, nullptr, nullptr);
  if (sym_diagonals) *sym_diagonals = h->nsym;
  if (sym_state) *sym_state = h->sym_state;
  if (offsets) *offsets = h->ntab;
  if (patterns) *patterns = h->npat;
  if (n) *n = h->nt;
  if (nnz) *nnz = h->nnz;
  if (spmv_kind) *spmv_kind = h->kind;
  if (spmv_grid) *spmv_grid = h->kind >= 0 ? h->grid[h->kind] : 0;
  if (stage_blocks) *stage_blocks = h->nblk[0];
  return TFG_OK;
}

int tfg_spmv_select(tfg_matrix* h, int kind) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (h->inner) return tfg_spmv_select(h->inner, kind);
  if (kind < -1 || kind >= tfg::kKinds) return fail(TFG_EINVAL, "unknown SpMV engine %d", kind);
  if (kind >= 0 && !kind_available(h, kind))
    return fail(TFG_ENOTSUP, "engine %d is not available on this matrix", kind);
  h->kind = kind;
  return TFG_OK;
}

int tfg_spmv_tune(tfg_matrix* h, const double* d_val, int* kind, float* us) {
  if (!h || !d_val) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner) {
    HIP_TRY(to_inner_val(h, d_val));
    return tfg_spmv_tune(h->inner, h->pval, kind, us);
  }
  int rc = tune_all(h, d_val);
  if (rc != TFG_OK) return rc;
  if (kind) *kind = h->kind;
  if (us)
    for (int k = 0; k < tfg::kKinds; ++k) us[k] = h->tune_us[k];
  return TFG_OK;
}

int tfg_solve_native(tfg_matrix* h, const char* solver, int solver_len, const char* prec,
                     int prec_len, const double* val, double* x, double* b, int miter,
                     int* niter, double tol, double* fin_res, const double* norm) {
  return tfg_solve_native_ex(h, solver, solver_len, prec, prec_len, val, x, b, miter, niter,
                             tol, fin_res, norm, 0, nullptr, nullptr);
}

int tfg_solve_native_ex(tfg_matrix* h, const char* solver, int solver_len, const char* prec,
                        int prec_len, const double* val, double* x, double* b, int miter,
                        int* niter, double tol, double* fin_res, const double* norm,
                        int device_ptrs, const tfg_options* opt, tfg_report* rep) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (!val || !x || !b || !niter || !fin_res) return fail(TFG_EINVAL, "null argument");
  // Solve_Native.f90:33-44: 'bicg' and 'cg' only.  The reference's "bicg"
  // applies A and M (not their transposes) to the shadow vectors and starts
  // them equal to the primal ones (Bicg.f90:129-133,143-195), so its iterates
  // are bit-identical to Cg's; it runs the CG recurrence here.
  // Beyond the reference: the textbook methods under their PETSc names
  // ('cgs', 'bcgs' / 'bicgstab'), and textbook BiCG for 'bicg' when
  // opt->textbook_bicg is set.
  if (h->inner && !is_icc(prec, prec_len)) {
    // internal ordering: permute in, solve the renumbered system, permute x
    // out.  ICC depends on the ordering, so it runs in the caller's order.
    HIP_TRY(hipSetDevice(h->device));
    hipStream_t S = h->inner->stream;
    const size_t n = (size_t)h->nt;
    const double *dval = val, *db = b;
    double* dx = x;
    if (!device_ptrs) {
      if (!h->val) HIP_TRY(hipMalloc(&h->val, sizeof(double) * (size_t)h->nnz));
      if (!h->b) HIP_TRY(hipMalloc(&h->b, sizeof(double) * n));
      HIP_TRY(hipMemcpyAsync(h->val, val, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice, S));
      HIP_TRY(hipMemcpyAsync(h->b, b, sizeof(double) * n, hipMemcpyHostToDevice, S));
      HIP_TRY(hipMemcpyAsync(h->x, x, sizeof(double) * n, hipMemcpyHostToDevice, S));
      dval = h->val;
      db = h->b;
      dx = h->x;
    }
    HIP_TRY(to_inner_val(h, dval));
    HIP_TRY(to_inner_vec(h, db, h->pb));
    HIP_TRY(to_inner_vec(h, dx, h->px));
    const int rc = tfg_solve_native_ex(h->inner, solver, solver_len, prec, prec_len, h->pval,
                                       h->px, h->pb, miter, niter, tol, fin_res, norm, 1, opt, rep);
    if (rc != TFG_OK && rc != TFG_EBREAKDOWN) return rc;
    HIP_TRY(from_inner_vec(h, h->px, dx));
    if (!device_ptrs) HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    return rc;
  }
  tfg_options o{};
  if (opt) o = *opt;
  const std::string sv = fstring(solver, solver_len);
  int method = tfg::NS_NONE;
  if (sv == "cgs") method = tfg::NS_CGS;
  else if (sv == "bcgs" || sv == "bicgstab") method = tfg::NS_BCGS;
  else if (sv == "bicg" && o.textbook_bicg) method = tfg::NS_BICG;
  else if (sv != "cg" && sv != "bicg")
    return fail(TFG_EINVAL, "Unknown native solver: %s", sv.c_str());
  // Prec_Form.f90:47-89: 'jacobi', 'icc', anything else = none.
  const std::string pv = fstring(prec, prec_len);
  const int pr = pv == "jacobi" ? tfg::PREC_JACOBI : (pv == "icc" ? tfg::PREC_ICC : tfg::PREC_NONE);
  if (pr == tfg::PREC_ICC && method != tfg::NS_NONE)
    return fail(TFG_ENOTSUP, "preconditioner 'icc' runs with cg / bicg only");

  if (o.criterion != TFG_CRIT_NRMS && o.criterion != TFG_CRIT_RELRES)
    return fail(TFG_EINVAL, "unknown criterion %d", o.criterion);
  if (o.precision != TFG_FP64 && o.precision != TFG_FP32_IR)
    return fail(TFG_EINVAL, "unknown precision %d", o.precision);
  if (o.path < TFG_PATH_AUTO || o.path > TFG_PATH_GRAPH)
    return fail(TFG_EINVAL, "unknown path %d", o.path);
  if (h->ni < h->nt && !h->allreduce)
    return fail(TFG_ENOTSUP, "buffer cells (ni=%d < nt=%d) need the tfg_matrix_set_comm hooks",
                h->ni, h->nt);
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->true_relres = -1.0;
  }
  HIP_TRY(hipSetDevice(h->device));
  const size_t n = (size_t)h->nt;
  const double* dval = val;
  const double* db = b;
  if (!device_ptrs) {
    if (!h->val) HIP_TRY(hipMalloc(&h->val, sizeof(double) * (size_t)h->nnz));
    if (!h->b) HIP_TRY(hipMalloc(&h->b, sizeof(double) * n));
    HIP_TRY(hipMemcpyAsync(h->val, val, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->b, b, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->x, x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    dval = h->val;
    db = h->b;
  } else {
    HIP_TRY(hipMemcpyAsync(h->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
  }

  int rc;
  if (h->allreduce) {
    if (method != tfg::NS_NONE || o.precision != TFG_FP64)
      return fail(TFG_ENOTSUP, "several ranks: FP64 cg / bicg with jacobi, icc or none");
    rc = tfg::solve_comm(h, pr, dval, db, miter, niter, tol, fin_res, norm, o, rep);
    if (rc != TFG_OK && rc != TFG_EBREAKDOWN) return rc;
  } else if (o.precision == TFG_FP32_IR && method != tfg::NS_NONE) {
    rc = tfg::solve_ir_ns(h, method, pr, dval, db, miter, niter, tol, fin_res, norm, o, rep);
    if (rc != TFG_OK && rc != TFG_EBREAKDOWN) return rc;
  } else if (o.precision == TFG_FP32_IR) {
    rc = tfg::solve_ir(h, pr, dval, db, miter, niter, tol, fin_res, norm, o, rep);
    if (rc != TFG_OK && rc != TFG_EBREAKDOWN) return rc;
  } else {
    { const int rt = ensure_tuned(h, dval); if (rt != TFG_OK) return rt; }
    tfg::Launch L = launch_of(h, dval, h->kind);
    tfg::CgArgs a = args_of(h, h->kind);
    a.val = dval;
    a.b = db;
    a.x = h->x;
    a.r = h->r;
    a.p = h->p;
    a.q = h->q;
    a.ad = h->ad;
    a.pt = partials_of(h, L);
    a.st = h->st;
    if (method == tfg::NS_CGS) a.pt.g_upd = L.grid_spmv;
    tfg::NsArgs ns{};
    if (method != tfg::NS_NONE) {
      const int rn = ensure_nonsym(h, method, dval, L);
      if (rn != TFG_OK) return rn;
      ns.valt = h->valt;
      ns.rt = h->rt;
      ns.v1 = h->nv[0];
      ns.v2 = h->nv[1];
      ns.v3 = h->nv[2];
      ns.v4 = h->nv[3];
    }

    State* s0 = &h->hst[0];
    memset(s0, 0, sizeof(State));
    s0->tol = tol;
    s0->criterion = o.criterion;
    s0->has_norm = norm ? 1 : 0;
    s0->norm = norm ? *norm : 0.0;
    s0->kcur = 1;
    HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof(State), hipMemcpyHostToDevice, h->stream));

    const bool timing = o.timing != 0;
    const int max_samples = 256;
    if (timing && h->ev_sample.empty()) {
      h->ev_sample.resize(2 * max_samples, nullptr);
      for (auto& e : h->ev_sample) HIP_TRY(hipEventCreate(&e));
    }
    std::vector<int> sample_iter;

    HIP_TRY(hipEventRecord(h->ev_t0, h->stream));
    // CG products by the symmetric engine when val is symmetric bit for bit
    if (h->kind == tfg::K_SYM_DIA && method == tfg::NS_NONE) {
      const int rs = sym_prepare(h, dval, &L.sym);
      if (rs != TFG_OK) return rs;
      a.sd = h->sym;
    }
    HIP_TRY(tfg::launch_diag_sum(L, a));
    const bool icc = pr == tfg::PREC_ICC;
    tfg::IccArgs ic{};
    if (icc) {
      const int ri = ensure_icc(h);
      if (ri != TFG_OK) return ri;
      ic = icc_args(h);
    }
    // ICC: the products and updates run unpreconditioned; z comes from the
    // sweeps (PREC_NONE kernels never read the Jacobi diagonal)
    const int pk = icc ? tfg::PREC_NONE : pr;
    HIP_TRY(tfg::launch_prologue(L, a, pk));
    if (icc) HIP_TRY(tfg::launch_icc_form(h->stream, ic, dval, h->st));
    // rho_1: (z, r0) for CG and BiCG, (r~0, r0) = (r0, r0) for CGS / BiCGSTAB
    HIP_TRY(tfg::launch_start(L, a, (method == tfg::NS_CGS || method == tfg::NS_BCGS) ? 5 : 6));
    if (method != tfg::NS_NONE) HIP_TRY(tfg::launch_ns_init(L, a, ns));

    // x update deferred into the next p update when no x range is needed
    const bool dx = method == tfg::NS_NONE && !icc && (o.criterion == TFG_CRIT_RELRES || norm);
    const int chunk = h->nnz > (1 << 22) ? 8 : 32;
    // Graph path: chunks of CG iterations replayed as one hipGraph; the
    // kernels take the iteration number from State::kcur (k = 0), which the
    // SpMV advances: three launches per iteration.  The executable graph is
    // kept per handle for the same operands.
    const bool graph = method == tfg::NS_NONE && !icc && !timing &&
                       (o.path == TFG_PATH_GRAPH || (o.path == TFG_PATH_AUTO && h->nnz < (1 << 21)));
    auto enqueue_dev_iterations = [&](int count, bool finish) -> int {
      for (int c = 0; c < count; ++c) {
        HIP_TRY(tfg::launch_pupdate(L, a, pr, 0, dx));
        HIP_TRY(tfg::launch_spmv_pq(L, a, 0));
        HIP_TRY(tfg::launch_update(L, a, pr, 0, dx));
      }
      if (finish) HIP_TRY(tfg::launch_finish(L, a, 0, 0));
      return TFG_OK;
    };
    if (graph) {
      const GraphKey key{dval, db, pr, dx ? 1 : 0, chunk, h->kind, L.sym ? 1 : 0};
      if (!h->cg_graph || !(h->cg_key == key)) {
        if (h->cg_graph) {
          HIP_TRY(hipGraphExecDestroy(h->cg_graph));
          h->cg_graph = nullptr;
        }
        hipGraph_t gr = nullptr;
        HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int rq = enqueue_dev_iterations(chunk, true);
        hipError_t ge = hipStreamEndCapture(h->stream, &gr);
        if (rq != TFG_OK) {
          if (gr) (void)hipGraphDestroy(gr);
          return rq;
        }
        if (ge == hipSuccess) ge = hipGraphInstantiate(&h->cg_graph, gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        HIP_TRY(ge);
        h->cg_key = key;
      }
    }
    int k = 1, slot = 0, prev = -1;
    bool stop = false;
    while (k <= miter && !stop) {
      const int kend = std::min(miter, k + chunk - 1);
      if (graph) {
        if (kend - k + 1 == chunk) {
          HIP_TRY(hipGraphLaunch(h->cg_graph, h->stream));
        } else {
          const int rq = enqueue_dev_iterations(kend - k + 1, true);
          if (rq != TFG_OK) return rq;
        }
        k = kend + 1;
      }
      for (; k <= kend; ++k) {
        const bool smp = timing && (k % 4 == 1) && (int)sample_iter.size() < max_samples;
        hipEvent_t e0 = smp ? h->ev_sample[2 * sample_iter.size()] : nullptr;
        hipEvent_t e1 = smp ? h->ev_sample[2 * sample_iter.size() + 1] : nullptr;
        if (method != tfg::NS_NONE) {
          HIP_TRY(tfg::launch_ns_iteration(L, a, ns, method, pr, k, e0, e1));
        } else if (icc) {
          if (k > 1) HIP_TRY(tfg::launch_icc_check(L, a, k));
          HIP_TRY(tfg::launch_icc_apply(h->stream, ic, dval, h->st, h->r, h->icc_z));
          HIP_TRY(tfg::launch_icc_rho(L, a, h->icc_z));
          HIP_TRY(tfg::launch_icc_pupdate(L, a, h->icc_z, k));
          if (smp) HIP_TRY(hipEventRecord(e0, h->stream));
          HIP_TRY(tfg::launch_spmv_pq(L, a, k));
          if (smp) HIP_TRY(hipEventRecord(e1, h->stream));
          HIP_TRY(tfg::launch_update(L, a, tfg::PREC_NONE, k));
        } else {
          HIP_TRY(tfg::launch_pupdate(L, a, pr, k, dx));
          if (smp) HIP_TRY(hipEventRecord(e0, h->stream));
          HIP_TRY(tfg::launch_spmv_pq(L, a, k));
          if (smp) HIP_TRY(hipEventRecord(e1, h->stream));
          HIP_TRY(tfg::launch_update(L, a, pr, k, dx));
        }
        if (smp) sample_iter.push_back(k);
      }
      if (!graph) HIP_TRY(tfg::launch_finish(L, a, kend, method == tfg::NS_BCGS ? 1 : 0));
      HIP_TRY(hipMemcpyAsync(&h->hst[1 + slot], h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipEventRecord(h->ev_poll[slot], h->stream));
      if (prev >= 0) {
        HIP_TRY(hipEventSynchronize(h->ev_poll[prev]));
        if (h->hst[1 + prev].done) stop = true;
      }
      prev = slot;
      slot ^= 1;
    }
    if (dx) HIP_TRY(tfg::launch_fold_x(L, a));
    HIP_TRY(hipEventRecord(h->ev_t1, h->stream));
    State* sf = &h->hst[3];
    HIP_TRY(hipMemcpyAsync(sf, h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *niter = sf->done ? sf->niter : std::max(miter, 0);
    *fin_res = sf->res;
    rc = sf->status;
    if (rep) {
      rep->status = rc;
      rep->iterations = *niter;
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
      rep->solve_ms = ms;
      double tot = 0.0;
      int cnt = 0;
      for (size_t s = 0; s < sample_iter.size(); ++s) {
        if (sample_iter[s] > *niter) break;  // launches after the stop return at once
        float e = 0.f;
        HIP_TRY(hipEventElapsedTime(&e, h->ev_sample[2 * s], h->ev_sample[2 * s + 1]));
        tot += e;
        ++cnt;
      }
      rep->spmv_ms = cnt ? tot / cnt : 0.0;
      rep->spmv_samples = cnt;
    }
    if (o.true_residual) {
      tfg::CgArgs t = a;
      t.val = dval;
      t.b = db;
      HIP_TRY(tfg::launch_true_residual(L, t));
      HIP_TRY(hipMemcpyAsync(sf, h->st, sizeof(State), hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      if (rep) rep->true_relres = sf->aux0;
    }
  }

  if (!device_ptrs)
    HIP_TRY(hipMemcpyAsync(x, h->x, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  else
    HIP_TRY(hipMemcpyAsync(x, h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (rc == TFG_EBREAKDOWN) return fail(TFG_EBREAKDOWN, "Krylov breakdown at iteration %d", *niter);
  return TFG_OK;
}

// The FP64 CG / reference-BiCG of tfg_solve_native_ex (device pointers,
// kernel path: the graph path gives the same bits) with its device work in
// the static scratch array instead of the handle's buffers; nothing is
// allocated (the pinned polling states and events are the handle's, made at
// create).  Layout of the scratch, each piece 256-byte aligned: r, p, q, the
// scaled diagonal (nt doubles each), the partial sums (kPartRows x pstride),
// the State, then with the symmetric engine its diagonals ((K + 1) x stride)
// and the asymmetry flag.  What must start at zero (the partials, the
// diagonals' gaps, the flag) is cleared by every call, since the array keeps
// the previous call's contents.
int tfg_solve_native_noalloc(tfg_matrix* h, const char* solver, int solver_len, const char* prec,
                             int prec_len, const double* d_val, double* d_x, const double* d_b,
                             int miter, int* niter, double tol, double* fin_res,
                             const double* norm, const tfg_options* opt, tfg_report* rep) {
  if (!h) return fail(TFG_EINVAL, "null handle");
  if (!d_val || !d_x || !d_b || !niter || !fin_res) return fail(TFG_EINVAL, "null argument");
  tfg_options o{};
  if (opt) o = *opt;
  const std::string sv = fstring(solver, solver_len);
  if (sv == "cgs" || sv == "bcgs" || sv == "bicgstab" || (sv == "bicg" && o.textbook_bicg))
    return fail(TFG_ENOTSUP, "the allocation-free solve runs the reference's cg / bicg only");
  if (sv != "cg" && sv != "bicg") return fail(TFG_EINVAL, "Unknown native solver: %s", sv.c_str());
  const std::string pv = fstring(prec, prec_len);
  const int pr = pv == "jacobi" ? tfg::PREC_JACOBI : (pv == "icc" ? tfg::PREC_ICC : tfg::PREC_NONE);
  if (pr == tfg::PREC_ICC)
    return fail(TFG_ENOTSUP, "the allocation-free solve takes jacobi or none, not icc");
  if (o.precision != TFG_FP64)
    return fail(TFG_ENOTSUP, "the allocation-free solve runs in FP64 only");
  if (o.criterion != TFG_CRIT_NRMS && o.criterion != TFG_CRIT_RELRES)
    return fail(TFG_EINVAL, "unknown criterion %d", o.criterion);
  if (o.path < TFG_PATH_AUTO || o.path > TFG_PATH_GRAPH)
    return fail(TFG_EINVAL, "unknown path %d", o.path);
  if (o.timing) return fail(TFG_ENOTSUP, "the allocation-free solve does not sample products");
  if (h->inner) return fail(TFG_ENOTSUP, "the allocation-free solve needs a matrix without "
                                         "internal ordering");
  if (h->ni < h->nt || h->allreduce)
    return fail(TFG_ENOTSUP, "the allocation-free solve needs a matrix without buffer cells");
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->true_relres = -1.0;
  }
  unsigned char* base = scratch_of(h->device);
  if (!base) return fail(TFG_EHIP, "the scratch array of device %d is not set up", h->device);

  // the handle's engine, else the one chosen without timing (not stored:
  // a later tfg_solve_native_ex may still time the engines)
  const int kind = h->kind >= 0 ? h->kind : default_kind(h);
  const bool symk = kind == tfg::K_SYM_DIA;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t n = (size_t)h->nt;
  const size_t vec = al(n * sizeof(double));
  const size_t partb = al(sizeof(double) * kPartRows * (size_t)h->pstride);
  const size_t stb = al(sizeof(State));
  const size_t sdb = symk ? al(sizeof(double) * (size_t)(h->nsym + 1) * (size_t)h->sdg.stride) : 0;
  const size_t flb = symk ? 256 : 0;
  const size_t need = 4 * vec + partb + stb + sdb + flb;
  if (need > (size_t)TFG_NOALLOC_MAX_BYTES)
    return fail(TFG_EINVAL,
                "the allocation-free solve of %d rows needs %zu bytes of scratch; it holds %lld "
                "(TFG_NOALLOC_MAX_BYTES)", h->nt, need, (long long)TFG_NOALLOC_MAX_BYTES);
  unsigned char* cur = base;
  auto take = [&cur](size_t b) { unsigned char* p = cur; cur += b; return p; };
  double* r = reinterpret_cast<double*>(take(vec));
  double* p = reinterpret_cast<double*>(take(vec));
  double* q = reinterpret_cast<double*>(take(vec));
  double* ad = reinterpret_cast<double*>(take(vec));
  double* part = reinterpret_cast<double*>(take(partb));
  State* st = reinterpret_cast<State*>(take(stb));
  double* sd = symk ? reinterpret_cast<double*>(take(sdb)) : nullptr;
  int* flag = symk ? reinterpret_cast<int*>(take(flb)) : nullptr;

  HIP_TRY(hipSetDevice(h->device));
  hipStream_t S = h->stream;
  // zeroed by kernels: a caller's stream may be under capture, where memset
  // nodes are avoided
  HIP_TRY(tfg::launch_zero16(S, part, (long long)(partb / 16)));
  tfg::Launch L = launch_of(h, d_val, kind);
  tfg::CgArgs a = args_of(h, kind);
  a.val = d_val;
  a.b = d_b;
  a.x = d_x;
  a.r = r;
  a.p = p;
  a.q = q;
  a.ad = ad;
  a.pt = partials_of(h, L, part);
  a.st = st;
  if (symk) {
    // the diagonals' absent entries and slot gaps must read zero
    HIP_TRY(tfg::launch_zero16(S, sd, (long long)(sdb / 16)));
    HIP_TRY(tfg::launch_zero16(S, flag, (long long)(flb / 16)));
    HIP_TRY(tfg::launch_sym_pack(S, grid_vec_of(h), shape_of(h, tfg::K_SYM_DIA), d_val, sd, flag));
    int asym = 1;
    HIP_TRY(hipMemcpyAsync(&asym, flag, sizeof(int), hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    L.sym = asym == 0;
    h->sym_state = L.sym ? 1 : 0;
    a.sd = sd;
  }

  State* s0 = &h->hst[0];
  memset(s0, 0, sizeof(State));
  s0->tol = tol;
  s0->criterion = o.criterion;
  s0->has_norm = norm ? 1 : 0;
  s0->norm = norm ? *norm : 0.0;
  s0->kcur = 1;
  HIP_TRY(hipMemcpyAsync(st, s0, sizeof(State), hipMemcpyHostToDevice, S));
  HIP_TRY(hipEventRecord(h->ev_t0, S));
  HIP_TRY(tfg::launch_diag_sum(L, a));
  HIP_TRY(tfg::launch_prologue(L, a, pr));
  HIP_TRY(tfg::launch_start(L, a, 6));
  const bool dx = o.criterion == TFG_CRIT_RELRES || norm;
  const int chunk = h->nnz > (1 << 22) ? 8 : 32;
  int k = 1, slot = 0, prev = -1;
  bool stop = false;
  while (k <= miter && !stop) {
    const int kend = std::min(miter, k + chunk - 1);
    for (; k <= kend; ++k) {
      HIP_TRY(tfg::launch_pupdate(L, a, pr, k, dx));
      HIP_TRY(tfg::launch_spmv_pq(L, a, k));
      HIP_TRY(tfg::launch_update(L, a, pr, k, dx));
    }
    HIP_TRY(tfg::launch_finish(L, a, kend, 0));
    HIP_TRY(hipMemcpyAsync(&h->hst[1 + slot], st, sizeof(State), hipMemcpyDeviceToHost, S));
    HIP_TRY(hipEventRecord(h->ev_poll[slot], S));
    if (prev >= 0) {
      HIP_TRY(hipEventSynchronize(h->ev_poll[prev]));
      if (h->hst[1 + prev].done) stop = true;
    }
    prev = slot;
    slot ^= 1;
  }
  if (dx) HIP_TRY(tfg::launch_fold_x(L, a));
  HIP_TRY(hipEventRecord(h->ev_t1, S));
  State* sf = &h->hst[3];
  HIP_TRY(hipMemcpyAsync(sf, st, sizeof(State), hipMemcpyDeviceToHost, S));
  HIP_TRY(hipStreamSynchronize(S));
  *niter = sf->done ? sf->niter : std::max(miter, 0);
  *fin_res = sf->res;
  const int rc = sf->status;
  if (rep) {
    rep->status = rc;
    rep->iterations = *niter;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
    rep->solve_ms = ms;
  }
  if (o.true_residual) {
    HIP_TRY(tfg::launch_true_residual(L, a));
    HIP_TRY(hipMemcpyAsync(sf, st, sizeof(State), hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    if (rep) rep->true_relres = sf->aux0;
  }
  if (rc == TFG_EBREAKDOWN) return fail(TFG_EBREAKDOWN, "Krylov breakdown at iteration %d", *niter);
  return TFG_OK;
}

int tfg_prec_apply(tfg_matrix* h, const char* prec, int prec_len, const double* d_val,
                   const double* d_r, double* d_z) {
  if (!h || !d_val || !d_r || !d_z) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner && !is_icc(prec, prec_len)) {
    HIP_TRY(to_inner_val(h, d_val));
    HIP_TRY(to_inner_vec(h, d_r, h->pb));
    const int rc = tfg_prec_apply(h->inner, prec, prec_len, h->pval, h->pb, h->pq);
    if (rc != TFG_OK) return rc;
    HIP_TRY(from_inner_vec(h, h->pq, d_z));
    return TFG_OK;
  }
  const std::string pv = fstring(prec, prec_len);
  if (pv == "icc") {
    const int ri = ensure_icc(h);
    if (ri != TFG_OK) return ri;
    const tfg::IccArgs ic = icc_args(h);
    HIP_TRY(tfg::launch_icc_form(h->stream, ic, d_val, nullptr));
    HIP_TRY(tfg::launch_icc_apply(h->stream, ic, d_val, nullptr, d_r, d_z));
  } else if (pv == "jacobi") {
    HIP_TRY(tfg::launch_jacobi_apply(h->stream, h->ni, h->row, d_val, d_r, d_z));
  } else {
    HIP_TRY(hipMemcpyAsync(d_z, d_r, sizeof(double) * (size_t)h->ni, hipMemcpyDeviceToDevice, h->stream));
  }
  return TFG_OK;
}

int tfg_spmv_fp32(tfg_matrix* h, const double* d_val, double fn, const float* d_x, float* d_y) {
  if (!h || !d_val || !d_x || !d_y) return fail(TFG_EINVAL, "null argument");
  if (h->inner) return fail(TFG_ENOTSUP, "FP32 product on an internally ordered matrix");
  if (h->ni < h->nt) return fail(TFG_ENOTSUP, "FP32 product with buffer cells");
  HIP_TRY(hipSetDevice(h->device));
  { const int rt = ensure_tuned(h, d_val); if (rt != TFG_OK) return rt; }
  for (float** v : {&h->ad32, &h->s32})
    if (!*v) HIP_TRY(hipMalloc(v, sizeof(float) * std::max<size_t>((size_t)h->nt, 4)));
  if (!h->s64) HIP_TRY(hipMalloc(&h->s64, sizeof(double) * std::max<size_t>((size_t)h->nt, 4)));
  bool sym = false;
  if (h->nsym > 0) {
    const int rs = sym_prepare(h, d_val, &sym);
    if (rs != TFG_OK) return rs;
  }
  if (sym) {
    const int r3 = ensure_sym32(h);
    if (r3 != TFG_OK) return r3;
    if (h->quad32 && (((uintptr_t)d_x | (uintptr_t)d_y) % 16))
      return fail(TFG_EINVAL, "FP32 product: x and y must be 16-byte aligned");
  } else if (!h->val32) {
    HIP_TRY(hipMalloc(&h->val32, sizeof(float) * (size_t)std::max(h->nnz, 4)));
  }
  tfg::Launch L32 = launch_of(h, h->val32, h->kind);
  if (sym) {
    L32.kind = tfg::K_SYM_DIA;
    L32.grid_spmv = h->grid32;
  } else {
    if (L32.kind != tfg::K_STAGE2048 && L32.kind != tfg::K_PIPE2048_CODED &&
        L32.kind != tfg::K_PIPE_PATTERN)
      L32.kind = tfg::K_PIPE2048;
    L32.grid_spmv = h->grid[L32.kind];
    L32.vec_val = h->nnz >= 4;
  }
  tfg::Launch L = launch_of(h, d_val, h->kind);
  tfg::CgArgs a = args_of(h, h->kind);
  a.val = d_val;
  a.pt = partials_of(h, L);
  a.st = h->st;
  tfg::IrArgs f{};
  f.n = h->ni;
  f.val32 = h->val32;
  f.sd32 = h->sd32;
  f.stride32 = h->stride32;
  f.sorder32 = h->sorder32;
  f.nsorder32 = h->nsorder32;
  f.quad = h->quad32;
  f.s64 = h->s64;
  f.s32 = h->s32;
  f.rd32 = h->ad32;
  State* s0 = &h->hst[0];
  memset(s0, 0, sizeof(State));
  s0->fn = fn;
  HIP_TRY(hipMemcpyAsync(h->st, s0, sizeof(State), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(tfg::launch_ir_convert(L, L32, a, f, h->nnz));
  tfg::SpmvShape m = shape_of(h, L32.kind == tfg::K_SYM_DIA ? h->kind : L32.kind);
  HIP_TRY(tfg::launch_spmv_lap32(L32, m, f, d_x, d_y));
  return TFG_OK;
}

int tfg_spmv(tfg_matrix* h, const double* d_val, const double* d_p, double* d_q) {
  if (!h || !d_val || !d_p || !d_q) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner) {
    HIP_TRY(to_inner_val(h, d_val));
    HIP_TRY(to_inner_vec(h, d_p, h->px));
    const int rc = tfg_spmv(h->inner, h->pval, h->px, h->pq);
    if (rc != TFG_OK) return rc;
    HIP_TRY(from_inner_vec(h, h->pq, d_q));
    return TFG_OK;
  }
  { const int rt = ensure_tuned(h, d_val); if (rt != TFG_OK) return rt; }
  tfg::Launch L = launch_of(h, d_val, h->kind);
  if (h->kind == tfg::K_SYM_DIA) {
    const int rs = sym_prepare(h, d_val, &L.sym);
    if (rs != TFG_OK) return rs;
  }
  HIP_TRY(tfg::launch_spmv_plain(L, shape_of(h, h->kind), d_val, d_p, d_q));
  return TFG_OK;
}

int tfg_residual(tfg_matrix* h, const double* d_val, const double* d_b, const double* d_x,
                 double* d_r) {
  if (!h || !d_val || !d_b || !d_x || !d_r) return fail(TFG_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner) {
    HIP_TRY(to_inner_val(h, d_val));
    HIP_TRY(to_inner_vec(h, d_b, h->pb));
    HIP_TRY(to_inner_vec(h, d_x, h->px));
    const int rc = tfg_residual(h->inner, h->pval, h->pb, h->px, h->pq);
    if (rc != TFG_OK) return rc;
    HIP_TRY(from_inner_vec(h, h->pq, d_r));
    return TFG_OK;
  }
  { const int rt = ensure_tuned(h, d_val); if (rt != TFG_OK) return rt; }
  tfg::Launch L = launch_of(h, d_val, h->kind);
  HIP_TRY(tfg::launch_residual_plain(L, shape_of(h, h->kind), d_val, d_b, d_x, d_r));
  return TFG_OK;
}

int tfg_spmv_bench(tfg_matrix* h, const double* d_val, const double* d_p, double* d_q, int reps,
                   double* ms_avg) {
  if (!h || !ms_avg || reps < 1) return fail(TFG_EINVAL, "bad argument");
  HIP_TRY(hipSetDevice(h->device));
  if (h->inner) {   // times the renumbered matrix's product alone
    HIP_TRY(to_inner_val(h, d_val));
    HIP_TRY(to_inner_vec(h, d_p, h->px));
    return tfg_spmv_bench(h->inner, h->pval, h->px, h->pq, reps, ms_avg);
  }
  { const int rt = ensure_tuned(h, d_val); if (rt != TFG_OK) return rt; }
  tfg::Launch L = launch_of(h, d_val, h->kind);
  if (h->kind == tfg::K_SYM_DIA) {
    const int rs = sym_prepare(h, d_val, &L.sym);
    if (rs != TFG_OK) return rs;
  }
  const tfg::SpmvShape m = shape_of(h, h->kind);
  HIP_TRY(hipEventRecord(h->ev_t0, h->stream));
  for (int i = 0; i < reps; ++i) HIP_TRY(tfg::launch_spmv_plain(L, m, d_val, d_p, d_q));
  HIP_TRY(hipEventRecord(h->ev_t1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev_t1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev_t0, h->ev_t1));
  *ms_avg = ms / reps;
  return TFG_OK;
}

int tfg_device_alloc(void** dptr, long long bytes) {
  if (!dptr || bytes <= 0) return fail(TFG_EINVAL, "bad allocation request");
  HIP_TRY(hipMalloc(dptr, (size_t)bytes));
  return TFG_OK;
}

int tfg_device_free(void* dptr) {
  HIP_TRY(hipFree(dptr));
  return TFG_OK;
}

int tfg_memcpy_h2d(void* dst, const void* src, long long bytes) {
  HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return TFG_OK;
}

int tfg_memcpy_d2h(void* dst, const void* src, long long bytes) {
  HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return TFG_OK;
}

int tfg_stream_copy(void* d_dst, const void* d_src, long long bytes, int reps, double* ms_avg) {
  return tfg_stream_bench(d_dst, d_src, bytes, 0, reps, ms_avg);
}

int tfg_stream_bench(void* d_dst, const void* d_src, long long bytes, int mode, int reps,
                     double* ms_avg) {
  if (mode < 0 || mode > 2) return fail(TFG_EINVAL, "unknown sweep mode %d", mode);
  if (!d_dst || (mode != 2 && !d_src) || bytes < 16 || (bytes % 16) || reps < 1 || !ms_avg)
    return fail(TFG_EINVAL, "bad sweep request");
  if (((uintptr_t)d_dst % 16) || ((uintptr_t)d_src % 16)) return fail(TFG_EINVAL, "sweeps need 16-byte alignment");
  int dev = 0, cus = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = std::max(1, cus * 4);
  if (mode == 1 && bytes < (long long)grid * 8)
    return fail(TFG_EINVAL, "the read sweep writes %d sums to d_dst: it needs %lld bytes", grid,
                (long long)grid * 8);
  auto once = [&]() -> hipError_t {
    if (mode == 0) return tfg::launch_copy16(nullptr, grid, d_src, d_dst, bytes / 16);
    if (mode == 1) return tfg::launch_read16(nullptr, grid, d_src, static_cast<double*>(d_dst), bytes / 16);
    return tfg::launch_fill16(nullptr, grid, d_dst, bytes / 16);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t r = hipEventCreate(&e0);
  if (r == hipSuccess) r = hipEventCreate(&e1);
  if (r == hipSuccess) r = once();   // warm-up
  if (r == hipSuccess) r = hipEventRecord(e0, nullptr);
  for (int k = 0; k < reps && r == hipSuccess; ++k) r = once();
  if (r == hipSuccess) r = hipEventRecord(e1, nullptr);
  if (r == hipSuccess) r = hipEventSynchronize(e1);
  float ms = 0.f;
  if (r == hipSuccess) r = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  HIP_TRY(r);
  *ms_avg = ms / reps;
  return TFG_OK;
}

int tfg_sync_bench(int grid, int reps, double* us_barrier, double* us_boundary) {
  if (grid < 1 || reps < 1 || !us_barrier || !us_boundary) return fail(TFG_EINVAL, "bad argument");
  int dev = 0, cus = 0, coop = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return fail(TFG_ENOTSUP, "no cooperative launches on this device");
  if (!tfg::grid_barrier_fit(&per_cu) || grid > per_cu * cus)
    return fail(TFG_EINVAL, "%d workgroups cannot all be resident (at most %d)", grid, per_cu * cus);
  double* sink = nullptr;
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  hipError_t r = hipMalloc(&sink, sizeof(double));
  for (hipEvent_t& x : e)
    if (r == hipSuccess) r = hipEventCreate(&x);
  if (r == hipSuccess) r = tfg::launch_grid_barriers(nullptr, grid, 1, sink);   // warm-up
  if (r == hipSuccess) r = hipEventRecord(e[0], nullptr);
  if (r == hipSuccess) r = tfg::launch_grid_barriers(nullptr, grid, reps, sink);
  if (r == hipSuccess) r = hipEventRecord(e[1], nullptr);
  if (r == hipSuccess) r = tfg::launch_boundary(nullptr, grid, sink, 0);
  if (r == hipSuccess) r = hipEventRecord(e[2], nullptr);
  for (int i = 0; i < reps && r == hipSuccess; ++i) r = tfg::launch_boundary(nullptr, grid, sink, i);
  if (r == hipSuccess) r = hipEventRecord(e[3], nullptr);
  if (r == hipSuccess) r = hipEventSynchronize(e[3]);
  float ms_b = 0.f, ms_k = 0.f;
  if (r == hipSuccess) r = hipEventElapsedTime(&ms_b, e[0], e[1]);
  if (r == hipSuccess) r = hipEventElapsedTime(&ms_k, e[2], e[3]);
  for (hipEvent_t& x : e)
    if (x) (void)hipEventDestroy(x);
  if (sink) (void)hipFree(sink);
  HIP_TRY(r);
  *us_barrier = 1e3 * ms_b / reps;
  *us_boundary = 1e3 * ms_k / reps;
  return TFG_OK;
}

int tfg_synchronize(tfg_matrix* h) {
  if (!h) HIP_TRY(hipDeviceSynchronize());
  else HIP_TRY(hipStreamSynchronize(h->stream));
  if (h && h->inner) HIP_TRY(hipStreamSynchronize(h->inner->stream));
  return TFG_OK;
}

}  // extern "C"
