// CSR SpMV, vector kernels and the conjugate-gradient solver on gfx950.
//
// SpMV ("row-stream"): a workgroup owns 256 consecutive rows.  Its threads
// stream the rows' value/column arrays with coalesced loads, form every
// product vals[j] * x[cols[j]] (rounded, as the reference's a*b) and park it
// in LDS; then thread t adds up row r0+t's products in stored column order.
// That is exactly matvec_std's arithmetic (SparseMatrix_functions.hpp:566-572)
// with HBM traffic of one pass over vals/cols.  No MFMA: the operator is sparse.
//
// CG: one HIP stream, device-resident scalars, fused kernels, and chunks of
// iterations captured in a HIP graph and replayed; every kernel first checks a
// device "done" flag so the reference's data-dependent loop test
// (cg_solve.hpp:152) stops the work without host round trips.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "mfe_common.hpp"

namespace mfe {

// ---------------------------------------------------------------- SpMV ------

// ZERO_Y: start each row sum at 0*y[row] (matvec_std's beta*ycoefs[row]); the
// fast CG path starts at +0.0 and does not read y.  DOT: also emit the
// workgroup partial of x.y (p.Ap in CG) at partials[tile].
// Every load whose address does not depend on the tile's column stream (row
// offsets, y[row], x[row]) is issued first, so no memory round trip sits
// between the LDS barrier and the row sums.
template <bool ZERO_Y, bool DOT>
__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(
    int n, const int* __restrict__ ro, const int* __restrict__ cols,
    const double* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ partials, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double prod[kSpmvMaxNnz];
  __shared__ double red[4];
  const int tile = (int)blockIdx.x;
  const int r0 = tile * kSpmvRows;
  const int r1 = min(r0 + kSpmvRows, n);
  const int t = (int)threadIdx.x;
  const int row = r0 + t;
  const bool valid = row < r1;
  const int rr = valid ? row : r1 - 1;
  const int ra = ro[rr], rb = ro[rr + 1];
  const double y0 = ZERO_Y ? y[rr] : 0.0;
  const double xr = DOT ? x[rr] : 0.0;
  const int s = ro[r0];
  const int cnt = ro[r1] - s;
  if (cnt > 0) {
    // Every column and value load of the tile before the first use (indices
    // clamped into the tile, so all loads are unconditional), then every x
    // gather, then the products: 27 loads in flight per lane.
    constexpr int K = kSpmvMaxNnz / kBlock;
    int c[K];
    double v[K], xv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = t + k * kBlock;
      const int j = s + (i < cnt ? i : 0);
      c[k] = cols[j];
      v[k] = vals[j];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) xv[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = t + k * kBlock;
      if (i < cnt) prod[i] = v[k] * xv[k];
    }
  }
  __syncthreads();
  double sum = 0.0;
  if (valid) {
    if (ZERO_Y) sum = 0.0 * y0;
    const int a = ra - s, b = rb - s;
    int j = a;
    for (; j + 4 <= b; j += 4) {  // LDS reads ahead, additions still in order
      const double p0 = prod[j], p1 = prod[j + 1], p2 = prod[j + 2], p3 = prod[j + 3];
      sum += p0;
      sum += p1;
      sum += p2;
      sum += p3;
    }
    for (; j < b; ++j) sum += prod[j];
    y[row] = sum;
  }
  if (DOT) {
    const double v0 = block_sum(valid ? sum * xr : 0.0, red);
    if (t == 0) partials[tile] = v0;
  }
}

// Fallback for matrices with a 256-row tile of more than 27*256 non-zeros:
// one thread per row, same per-row order, so the same bits.
template <bool ZERO_Y, bool DOT>
__global__ __launch_bounds__(kBlock) void spmv_rowwise_kernel(
    int n, const int* __restrict__ ro, const int* __restrict__ cols,
    const double* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ partials, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[4];
  const int row = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double sum = 0.0;
  if (row < n) {
    if (ZERO_Y) sum = 0.0 * y[row];
    for (int j = ro[row]; j < ro[row + 1]; ++j) sum += vals[j] * x[cols[j]];
    y[row] = sum;
  }
  if (DOT) {
    const double v0 = block_sum(row < n ? sum * x[row] : 0.0, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v0;
  }
}

// Launch y = A x on stream; *tiles_out = number of partials written (DOT).
template <bool ZERO_Y, bool DOT>
static int launch_spmv(const mfe_csr* A, const double* x, double* y, double* partials,
                       const int* done, hipStream_t stream, int* tiles_out) {
  const int n = (int)A->nrows;
  if (n == 0) {
    if (tiles_out) *tiles_out = 0;
    return 0;
  }
  const int tiles = (int)cdiv(n, kSpmvRows);
  if (A->max_block_nnz >= 0 && A->max_block_nnz <= kSpmvMaxNnz) {
    spmv_stream_kernel<ZERO_Y, DOT><<<tiles, kBlock, 0, stream>>>(
        n, A->row_offsets, A->cols, A->vals, x, y, partials, done);
  } else {
    spmv_rowwise_kernel<ZERO_Y, DOT><<<tiles, kBlock, 0, stream>>>(
        n, A->row_offsets, A->cols, A->vals, x, y, partials, done);
  }
  if (tiles_out) *tiles_out = tiles;
  return (int)hipGetLastError();
}

// -------------------------------------------------------------- vectors -----

__global__ __launch_bounds__(kBlock) void waxpby_kernel(int64_t n, double alpha,
                                                        const double* x, double beta,
                                                        const double* y, double* w) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    w[i] = alpha * x[i] + beta * y[i];
}

// Per-thread partial sums over a fixed grid-stride pattern, then the fixed
// wave/workgroup tree: deterministic for a given n.
__device__ __forceinline__ double grid_dot(int64_t n, const double* __restrict__ x,
                                           const double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += x[i] * y[i];
  return acc;
}

__global__ __launch_bounds__(kBlock) void dot_partial_kernel(int64_t n,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             double* __restrict__ partials) {
  __shared__ double red[4];
  const double acc = block_sum(grid_dot(n, x, y), red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Workgroup sum of p[0..np): each thread adds its strided values in index
// order (eight loads in flight per step), then the fixed workgroup tree.
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int np,
                                               double* red) {
  double acc = 0.0;
  int i = (int)threadIdx.x;
  for (; i + 7 * kBlock < np; i += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; i < np; i += kBlock) acc += p[i];
  return block_sum(acc, red);
}

// First stage of a two-stage reduction of nvec arrays of np partials: block j
// sums partials[v*np + j*kStage .. +kStage) into out[v*gridDim.x + j].
constexpr int kStage = 8 * kBlock;
__global__ __launch_bounds__(kBlock) void reduce_stage_kernel(const double* __restrict__ partials,
                                                              int np, int nvec,
                                                              double* __restrict__ out,
                                                              const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[4];
  const int j = (int)blockIdx.x;
  const int lo = j * kStage;
  const int m = min(kStage, np - lo);
  for (int v = 0; v < nvec; ++v) {
    const double s = sum_partials(partials + (int64_t)v * np + lo, m, red);
    if (threadIdx.x == 0) out[(int64_t)v * gridDim.x + j] = s;
  }
}

__global__ __launch_bounds__(kBlock) void finalize_kernel(const double* __restrict__ partials,
                                                          int np, int nvec,
                                                          double* __restrict__ out) {
  __shared__ double red[4];
  for (int v = 0; v < nvec; ++v) {
    const double s = sum_partials(partials + (int64_t)v * np, np, red);
    if (threadIdx.x == 0) out[v] = s;
  }
}

// Sequential dot, the reference's left-to-right order (Vector_functions.hpp:
// 216-218): the workgroup stages products x[i]*y[i] into LDS in chunks and
// thread 0 adds them in index order.  One workgroup; for replay/parity use.
constexpr int kSeqChunk = 8 * kBlock;
__device__ double dot_sequential(int64_t n, const double* x, const double* y, double* buf) {
  double result = 0;
  for (int64_t base = 0; base < n; base += kSeqChunk) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t i = base + threadIdx.x + k * kBlock;
      if (i < n) buf[threadIdx.x + k * kBlock] = x[i] * y[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int m = (int)((n - base) < kSeqChunk ? (n - base) : kSeqChunk);
      for (int j = 0; j < m; ++j) result += buf[j];
    }
    __syncthreads();
  }
  return result;  // valid in thread 0
}

__global__ __launch_bounds__(kBlock) void dot_seq_kernel(int64_t n, const double* x,
                                                         const double* y, double* out) {
  __shared__ double buf[kSeqChunk];
  const double r = dot_sequential(n, x, y, buf);
  if (threadIdx.x == 0) *out = r;
}

// ------------------------------------------------------------------- CG -----
//
// Per iteration k (fast mode), all on one stream:
//   cg_p       x += alpha_{k-1} p_{k-1} (deferred, see below); p = 1*r + beta*y
//   spmv       Ap = A p, workgroup partials of p.Ap
//   reduce     first stage of the p.Ap reduction
//   alpha      p.Ap; breakdown test; alpha = rtrans / p.Ap
//   norms(+fin) only when p.Ap < brkdown_tol: ||Ap||, ||p|| for breakdown()
//   update     r = 1*r + (-alpha)*Ap, workgroup partials of r.r
//   next       end of k; loop test; rtrans, beta, normr of k+1
// The reference updates x right after r (cg_solve.hpp:211-212).  Here the
// update x = 1*x + alpha*p of iteration k is applied by the next cg_p, which
// reads p anyway (or by cg_finish after the last iteration): the same
// operation on the same operands, so the same bits, one vector pass fewer.

struct CgState {
  double rtrans, oldrtrans, pAp, alpha, beta, normr, brkdown_tol, tol;
  double alpha_prev;  // alpha of the iteration whose x update is pending
  int k, max_iter, done, status, num_iters, need_norms, x_pending, pad;
};

// Prologue scalar step: rtrans = r0.r0, normr, and the loop test for k = 1.
__device__ void cg_begin_first(CgState* st, double rr) {
  st->rtrans = rr;
  st->normr = sqrt(rr);
  st->beta = 0.0;
  st->k = 1;
  if (!(1 <= st->max_iter && st->normr > st->tol)) st->done = 1;
}

// End of iteration k (r updated): its x update becomes pending, num_iters = k,
// then the reference's loop test with the normr of iteration k, and rtrans,
// beta, normr of iteration k+1 (cg_solve.hpp:152-165).
__device__ void cg_next(CgState* st, double rr) {
  st->x_pending = 1;
  st->alpha_prev = st->alpha;
  st->num_iters = st->k;
  const int k1 = st->k + 1;
  if (!(k1 <= st->max_iter && st->normr > st->tol)) {
    st->done = 1;
    return;
  }
  st->oldrtrans = st->rtrans;
  st->rtrans = rr;
  st->beta = st->rtrans / st->oldrtrans;
  st->normr = sqrt(st->rtrans);
  st->k = k1;
}

__device__ double* hist_row(const CgState* st, double* hist) {
  return hist + 4 * (int64_t)(st->k - 1);
}

// Record p.Ap and, unless the breakdown test needs the norms, alpha.
// Returns 1 when alpha was set.  (cg_solve.hpp:186-201)
__device__ int cg_alpha_or_defer(CgState* st, double pAp, double* hist) {
  st->x_pending = 0;  // cg_p of this iteration has applied it
  st->pAp = pAp;
  double* h = hist_row(st, hist);
  h[0] = st->rtrans;
  h[1] = pAp;
  h[2] = 0.0;
  h[3] = st->beta;
  if (pAp < st->brkdown_tol) {
    if (pAp < 0) {
      st->status = MFE_BREAKDOWN;
      st->done = 1;
      return 0;
    }
    st->need_norms = 1;
    return 0;
  }
  st->alpha = st->rtrans / pAp;
  h[2] = st->alpha;
  return 1;
}

// breakdown() (cg_solve.hpp:49-67) with v = Ap, w = p, then alpha.
__device__ void cg_alpha_after_norms(CgState* st, double ApAp, double pp, double* hist) {
  st->need_norms = 0;
  const double vnorm = sqrt(ApAp), wnorm = sqrt(pp);
  if (fabs(st->pAp) <= 100 * vnorm * wnorm * DBL_EPSILON) {
    st->status = MFE_BREAKDOWN;
    st->done = 1;
    return;
  }
  st->brkdown_tol = 0.1 * st->pAp;
  st->alpha = st->rtrans / st->pAp;
  hist_row(st, hist)[2] = st->alpha;
}

__global__ void cg_init_kernel(CgState* st, int max_iter, double tol) {
  CgState s;
  memset(&s, 0, sizeof(s));
  s.max_iter = max_iter;
  s.tol = tol;
  s.brkdown_tol = DBL_EPSILON;
  s.k = 1;
  *st = s;
}

// Fast mode scalar kernels (one workgroup each) ------------------------------
__global__ __launch_bounds__(kBlock) void cg_first_fast_kernel(CgState* st,
                                                               const double* partials, int np) {
  __shared__ double red[4];
  const double rr = sum_partials(partials, np, red);
  if (threadIdx.x == 0) cg_begin_first(st, rr);
}

__global__ __launch_bounds__(kBlock) void cg_alpha_fast_kernel(CgState* st,
                                                               const double* partials, int np,
                                                               double* hist) {
  if (st->done) return;
  __shared__ double red[4];
  const double pAp = sum_partials(partials, np, red);
  if (threadIdx.x == 0) cg_alpha_or_defer(st, pAp, hist);
}

// ||Ap||^2 and ||p||^2 partials, only when the breakdown test needs them.
__global__ __launch_bounds__(kBlock) void cg_norms_kernel(int64_t n, const CgState* st,
                                                          const double* __restrict__ Ap,
                                                          const double* __restrict__ p,
                                                          double* __restrict__ partials) {
  if (st->done || !st->need_norms) return;
  __shared__ double red[4];
  const double a = block_sum(grid_dot(n, Ap, Ap), red);
  const double b = block_sum(grid_dot(n, p, p), red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[gridDim.x + blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kBlock) void cg_norms_fin_kernel(CgState* st, const double* partials,
                                                              int np, double* hist) {
  if (st->done || !st->need_norms) return;
  __shared__ double red[4];
  const double ApAp = sum_partials(partials, np, red);
  const double pp = sum_partials(partials + np, np, red);
  if (threadIdx.x == 0) cg_alpha_after_norms(st, ApAp, pp, hist);
}

__global__ __launch_bounds__(kBlock) void cg_next_fast_kernel(CgState* st,
                                                              const double* partials, int np) {
  if (st->done) return;
  __shared__ double red[4];
  const double rr = sum_partials(partials, np, red);
  if (threadIdx.x == 0) cg_next(st, rr);
}

// Replay mode scalar kernels (sequential dots) ---------------------------------
__global__ __launch_bounds__(kBlock) void cg_first_seq_kernel(CgState* st, int64_t n,
                                                              const double* r) {
  __shared__ double buf[kSeqChunk];
  const double rr = dot_sequential(n, r, r, buf);
  if (threadIdx.x == 0) cg_begin_first(st, rr);
}

// End of iteration k, the loop test for k+1 and, only if it passes, the
// sequential r.r of iteration k+1.
__global__ __launch_bounds__(kBlock) void cg_next_seq_kernel(CgState* st, int64_t n,
                                                             const double* r) {
  __shared__ double buf[kSeqChunk];
  __shared__ int go;
  if (threadIdx.x == 0) {
    go = 0;
    if (!st->done) {
      st->x_pending = 1;
      st->alpha_prev = st->alpha;
      st->num_iters = st->k;
      const int k1 = st->k + 1;
      if (!(k1 <= st->max_iter && st->normr > st->tol)) st->done = 1;
      else go = 1;
    }
  }
  __syncthreads();
  if (!go) return;
  const double rr = dot_sequential(n, r, r, buf);
  if (threadIdx.x == 0) {
    st->oldrtrans = st->rtrans;
    st->rtrans = rr;
    st->beta = st->rtrans / st->oldrtrans;
    st->normr = sqrt(st->rtrans);
    st->k = st->k + 1;
  }
}

__global__ __launch_bounds__(kBlock) void cg_alpha_seq_kernel(CgState* st, int64_t n,
                                                              const double* Ap, const double* p,
                                                              double* hist) {
  __shared__ double buf[kSeqChunk];
  __shared__ int need;
  if (st->done) return;
  const double pAp = dot_sequential(n, Ap, p, buf);
  if (threadIdx.x == 0) {
    cg_alpha_or_defer(st, pAp, hist);
    need = st->need_norms;
  }
  __syncthreads();
  if (!need) return;
  const double ApAp = dot_sequential(n, Ap, Ap, buf);
  const double pp = dot_sequential(n, p, p, buf);
  if (threadIdx.x == 0) cg_alpha_after_norms(st, ApAp, pp, hist);
}

// x = 1*x + alpha_prev*p if an x update is pending; then p = 1*r + beta*y
// (y = r for k = 1, y = p afterwards; cg_solve.hpp:153-161, 211).
__global__ __launch_bounds__(kBlock) void cg_p_kernel(int64_t n, const CgState* st,
                                                      const double* r, const double* y,
                                                      double* p, double* x) {
  if (st->done) return;
  const double beta = st->beta;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (st->x_pending) {
    const double a = st->alpha_prev;
    for (int64_t i = i0; i < n; i += stride) {
      const double pi = p[i];
      x[i] = 1.0 * x[i] + a * pi;
      p[i] = 1.0 * r[i] + beta * (y == p ? pi : y[i]);
    }
  } else {
    for (int64_t i = i0; i < n; i += stride) p[i] = 1.0 * r[i] + beta * y[i];
  }
}

// After the last iteration: apply the pending x update.
__global__ __launch_bounds__(kBlock) void cg_finish_kernel(int64_t n, const CgState* st,
                                                           const double* __restrict__ p,
                                                           double* __restrict__ x) {
  if (!st->x_pending) return;
  const double a = st->alpha_prev;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = 1.0 * x[i] + a * p[i];
}

// r = 1*r + (-alpha)*Ap (cg_solve.hpp:212); with DOT, partials of the new r.r.
template <bool DOT>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(int64_t n, const CgState* st,
                                                           double* __restrict__ r,
                                                           const double* __restrict__ Ap,
                                                           double* __restrict__ partials) {
  if (st->done) return;
  __shared__ double red[4];
  const double malpha = -st->alpha;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ri = 1.0 * r[i] + malpha * Ap[i];
    r[i] = ri;
    if (DOT) acc += ri * ri;
  }
  if (DOT) {
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
  }
}

struct CgWorkspace {
  int64_t n = -1;
  double *r = nullptr, *p = nullptr, *Ap = nullptr;
  double* partials = nullptr;
  int64_t partials_cap = 0;
  double* red2 = nullptr;  // second-stage partials (ceil(tiles / kStage))
  int64_t red2_cap = 0;
  CgState* st = nullptr;
  CgState* h_st = nullptr;  // pinned
  double* hist = nullptr;
  int64_t hist_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // What the captured graph was built for.
  const void *kA = nullptr, *kb = nullptr, *kx = nullptr;
  const double* khist = nullptr;
  unsigned kflags = 0;
  int kchunk = 0;
  int64_t knnz = -1;
};

constexpr int kGraphChunk = 25;  // CG iterations per captured graph

static void ws_drop_graph(CgWorkspace* w) {
  if (w->exec) hipGraphExecDestroy(w->exec);
  if (w->graph) hipGraphDestroy(w->graph);
  w->exec = nullptr;
  w->graph = nullptr;
  w->kA = w->kb = w->kx = nullptr;
}

static void ws_free(CgWorkspace* w) {
  ws_drop_graph(w);
  hipFree(w->r);
  hipFree(w->p);
  hipFree(w->Ap);
  hipFree(w->partials);
  hipFree(w->red2);
  hipFree(w->st);
  hipFree(w->hist);
  if (w->h_st) hipHostFree(w->h_st);
  delete w;
}

static int ws_prepare(mfe_ctx* ctx, int64_t n, int64_t tiles, int max_iter, CgWorkspace** out) {
  CgWorkspace* w = ctx->cg;
  if (!w) {
    w = new CgWorkspace();
    ctx->cg = w;
    MFE_CHECK(ctx, hipMalloc(&w->st, sizeof(CgState)));
    MFE_CHECK(ctx, hipHostMalloc(&w->h_st, sizeof(CgState), hipHostMallocDefault));
  }
  if (w->n != n) {
    ws_drop_graph(w);
    hipFree(w->r);
    hipFree(w->p);
    hipFree(w->Ap);
    w->r = w->p = w->Ap = nullptr;
    w->n = -1;
    const size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
    MFE_CHECK(ctx, hipMalloc(&w->r, bytes));
    MFE_CHECK(ctx, hipMalloc(&w->p, bytes));
    MFE_CHECK(ctx, hipMalloc(&w->Ap, bytes));
    w->n = n;
  }
  const int64_t need = 3 * (tiles > kVecBlocks ? tiles : kVecBlocks);
  if (w->partials_cap < need) {
    ws_drop_graph(w);
    hipFree(w->partials);
    w->partials = nullptr;
    w->partials_cap = 0;
    MFE_CHECK(ctx, hipMalloc(&w->partials, sizeof(double) * (size_t)need));
    w->partials_cap = need;
  }
  const int64_t need2 = cdiv(tiles, kStage) > 0 ? cdiv(tiles, kStage) : 1;
  if (w->red2_cap < need2) {
    ws_drop_graph(w);
    hipFree(w->red2);
    w->red2 = nullptr;
    w->red2_cap = 0;
    MFE_CHECK(ctx, hipMalloc(&w->red2, sizeof(double) * (size_t)need2));
    w->red2_cap = need2;
  }
  const int64_t hneed = 4 * (int64_t)(max_iter > 0 ? max_iter : 1);
  if (w->hist_cap < hneed) {
    ws_drop_graph(w);
    hipFree(w->hist);
    w->hist = nullptr;
    w->hist_cap = 0;
    MFE_CHECK(ctx, hipMalloc(&w->hist, sizeof(double) * (size_t)hneed));
    w->hist_cap = hneed;
  }
  *out = w;
  return MFE_OK;
}

static unsigned vec_grid(int64_t n) {
  const int64_t b = cdiv(n, kBlock);
  return (unsigned)(b < kVecBlocks ? (b > 0 ? b : 1) : kVecBlocks);
}

// One CG iteration body (k = 1 when y_for_p == r, k >= 2 when y_for_p == p).
static hipError_t enqueue_iteration(const mfe_csr* A, CgWorkspace* w, double* x,
                                    const double* y_for_p, bool replay, hipStream_t s) {
  const int64_t n = w->n;
  const unsigned vg = vec_grid(n);
  cg_p_kernel<<<vg, kBlock, 0, s>>>(n, w->st, w->r, y_for_p, w->p, x);
  int tiles = 0;
  const int* done = &w->st->done;
  if (replay) {
    launch_spmv<true, false>(A, w->p, w->Ap, nullptr, done, s, &tiles);
    cg_alpha_seq_kernel<<<1, kBlock, 0, s>>>(w->st, n, w->Ap, w->p, w->hist);
    cg_update_kernel<false><<<vg, kBlock, 0, s>>>(n, w->st, w->r, w->Ap, nullptr);
    cg_next_seq_kernel<<<1, kBlock, 0, s>>>(w->st, n, w->r);
  } else {
    launch_spmv<false, true>(A, w->p, w->Ap, w->partials, done, s, &tiles);
    const int nb2 = (int)cdiv(tiles, kStage);
    reduce_stage_kernel<<<nb2, kBlock, 0, s>>>(w->partials, tiles, 1, w->red2, done);
    cg_alpha_fast_kernel<<<1, kBlock, 0, s>>>(w->st, w->red2, nb2, w->hist);
    cg_norms_kernel<<<vg, kBlock, 0, s>>>(n, w->st, w->Ap, w->p, w->partials);
    cg_norms_fin_kernel<<<1, kBlock, 0, s>>>(w->st, w->partials, (int)vg, w->hist);
    cg_update_kernel<true><<<vg, kBlock, 0, s>>>(n, w->st, w->r, w->Ap, w->partials);
    cg_next_fast_kernel<<<1, kBlock, 0, s>>>(w->st, w->partials, (int)vg);
  }
  return hipGetLastError();
}

}  // namespace mfe

using namespace mfe;

extern "C" {

void mfe_cg_workspace_free(mfe_ctx* ctx) {
  if (ctx && ctx->cg) {
    ws_free(ctx->cg);
    ctx->cg = nullptr;
  }
}

static int check_vec(mfe_ctx* ctx, const mfe_vec* v, int64_t n, const char* what) {
  if (!v || v->n != n) return fail(ctx, MFE_BAD_ARG, std::string(what) + ": length mismatch");
  return MFE_OK;
}

static int ensure_partials(mfe_ctx* ctx, int64_t count) {
  if ((int64_t)ctx->partials_cap >= count) return MFE_OK;
  hipFree(ctx->d_partials);
  ctx->d_partials = nullptr;
  ctx->partials_cap = 0;
  MFE_CHECK(ctx, hipMalloc(&ctx->d_partials, sizeof(double) * (size_t)count));
  ctx->partials_cap = (size_t)count;
  return MFE_OK;
}

int mfe_matvec(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* x, mfe_vec* y) {
  if (!ctx || !A) return MFE_BAD_ARG;
  int st;
  if ((st = check_vec(ctx, x, A->nrows, "matvec x")) || (st = check_vec(ctx, y, A->nrows, "matvec y")))
    return st;
  int e = launch_spmv<true, false>(A, x->d, y->d, nullptr, nullptr, ctx->stream, nullptr);
  if (e) return hip_fail(ctx, (hipError_t)e, "spmv");
  return MFE_OK;
}

int mfe_matvec_and_dot(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* x, mfe_vec* y,
                       double* result) {
  if (!ctx || !A || !result) return MFE_BAD_ARG;
  int st;
  if ((st = check_vec(ctx, x, A->nrows, "matvec x")) || (st = check_vec(ctx, y, A->nrows, "matvec y")))
    return st;
  const int64_t tiles = cdiv(A->nrows, kSpmvRows);
  if ((st = ensure_partials(ctx, tiles > 0 ? tiles : 1))) return st;
  int nt = 0;
  int e = launch_spmv<true, true>(A, x->d, y->d, ctx->d_partials, nullptr, ctx->stream, &nt);
  if (e) return hip_fail(ctx, (hipError_t)e, "spmv_dot");
  if (nt == 0) {
    *result = 0;
    return MFE_OK;
  }
  finalize_kernel<<<1, kBlock, 0, ctx->stream>>>(ctx->d_partials, nt, 1, ctx->d_scalars);
  MFE_LAUNCHED(ctx, "finalize_kernel");
  MFE_CHECK(ctx, hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  MFE_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *result = ctx->h_scalars[0];
  return MFE_OK;
}

int mfe_time_matvec(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* x, mfe_vec* y, int variant,
                    int reps, double* avg_ms) {
  if (!ctx || !A || !avg_ms || reps <= 0 || variant < 0 || variant > 1) return MFE_BAD_ARG;
  int st;
  if ((st = check_vec(ctx, x, A->nrows, "matvec x")) || (st = check_vec(ctx, y, A->nrows, "matvec y")))
    return st;
  const int64_t tiles = cdiv(A->nrows, kSpmvRows);
  if ((st = ensure_partials(ctx, tiles > 0 ? tiles : 1))) return st;
  MFE_CHECK(ctx, hipEventRecord(ctx->t_start, ctx->stream));
  for (int i = 0; i < reps; ++i) {
    int e = variant == 0
                ? launch_spmv<true, false>(A, x->d, y->d, nullptr, nullptr, ctx->stream, nullptr)
                : launch_spmv<false, true>(A, x->d, y->d, ctx->d_partials, nullptr, ctx->stream,
                                           nullptr);
    if (e) return hip_fail(ctx, (hipError_t)e, "spmv");
  }
  MFE_CHECK(ctx, hipEventRecord(ctx->t_stop, ctx->stream));
  MFE_CHECK(ctx, hipEventSynchronize(ctx->t_stop));
  float ms = 0;
  MFE_CHECK(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
  *avg_ms = (double)ms / reps;
  return MFE_OK;
}

int mfe_dot(mfe_ctx* ctx, const mfe_vec* x, const mfe_vec* y, double* result) {
  if (!ctx || !x || !y || !result) return MFE_BAD_ARG;
  if (y->n < x->n) return fail(ctx, MFE_BAD_ARG, "dot: y shorter than x");
  const int64_t n = x->n;
  const unsigned g = vec_grid(n);
  int st;
  if ((st = ensure_partials(ctx, g))) return st;
  dot_partial_kernel<<<g, kBlock, 0, ctx->stream>>>(n, x->d, y->d, ctx->d_partials);
  MFE_LAUNCHED(ctx, "dot_partial_kernel");
  finalize_kernel<<<1, kBlock, 0, ctx->stream>>>(ctx->d_partials, (int)g, 1, ctx->d_scalars);
  MFE_LAUNCHED(ctx, "finalize_kernel");
  MFE_CHECK(ctx, hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  MFE_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *result = ctx->h_scalars[0];
  return MFE_OK;
}

int mfe_dot_replay(mfe_ctx* ctx, const mfe_vec* x, const mfe_vec* y, double* result) {
  if (!ctx || !x || !y || !result) return MFE_BAD_ARG;
  if (y->n < x->n) return fail(ctx, MFE_BAD_ARG, "dot: y shorter than x");
  dot_seq_kernel<<<1, kBlock, 0, ctx->stream>>>(x->n, x->d, y->d, ctx->d_scalars);
  MFE_LAUNCHED(ctx, "dot_seq_kernel");
  MFE_CHECK(ctx, hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  MFE_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *result = ctx->h_scalars[0];
  return MFE_OK;
}

int mfe_waxpby(mfe_ctx* ctx, double alpha, const mfe_vec* x, double beta, const mfe_vec* y,
               mfe_vec* w) {
  if (!ctx || !x || !y || !w) return MFE_BAD_ARG;
  if (y->n < x->n || w->n < x->n) return fail(ctx, MFE_BAD_ARG, "waxpby: y and w must be at least as long as x");
  if (x->n == 0) return MFE_OK;
  waxpby_kernel<<<vec_grid(x->n), kBlock, 0, ctx->stream>>>(x->n, alpha, x->d, beta, y->d, w->d);
  MFE_LAUNCHED(ctx, "waxpby_kernel");
  return MFE_OK;
}

int mfe_cg_solve(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* b, mfe_vec* x, int max_iter,
                 double tolerance, unsigned flags, int* num_iters, double* normr,
                 double* times, double* history) {
  if (!ctx || !A || !num_iters || !normr) return MFE_BAD_ARG;
  if (!A->has_local_indices)
    return fail(ctx, MFE_NOT_LOCAL,
                "cg_solve: A.has_local_indices is false; call make_local_matrix(A) first");
  const int64_t n = A->nrows;
  int st;
  if ((st = check_vec(ctx, b, n, "cg b")) || (st = check_vec(ctx, x, n, "cg x"))) return st;
  MFE_CHECK(ctx, hipSetDevice(ctx->device));
  const bool replay = (flags & MFE_CG_REPLAY) != 0;
  const bool eager = (flags & MFE_CG_EAGER) != 0;
  hipStream_t s = ctx->stream;

  double tol = tolerance;
  if (flags & MFE_CG_TOL_RELATIVE) {
    double bb = 0;
    if ((st = mfe_dot(ctx, b, b, &bb))) return st;
    tol = tolerance * std::sqrt(bb);
  }

  const int64_t tiles = cdiv(n, kSpmvRows);
  CgWorkspace* w = nullptr;
  if ((st = ws_prepare(ctx, n, tiles, max_iter, &w))) return st;

  MFE_CHECK(ctx, hipEventRecord(ctx->t_start, s));
  cg_init_kernel<<<1, 1, 0, s>>>(w->st, max_iter, tol);
  MFE_LAUNCHED(ctx, "cg_init_kernel");
  *num_iters = 0;
  if (n > 0) {
    const unsigned vg = vec_grid(n);
    // p = 1*x + 0*x ; Ap = A p ; r = 1*b + (-1)*Ap ; rtrans = r.r  (cg_solve.hpp:124-136)
    waxpby_kernel<<<vg, kBlock, 0, s>>>(n, 1.0, x->d, 0.0, x->d, w->p);
    // The reference's Ap starts as a zero vector, so its 0*Ap[row] is +0.0:
    // start the sums at +0.0 instead of reading the workspace.
    int e = launch_spmv<false, false>(A, w->p, w->Ap, nullptr, nullptr, s, nullptr);
    if (e) return hip_fail(ctx, (hipError_t)e, "cg prologue spmv");
    waxpby_kernel<<<vg, kBlock, 0, s>>>(n, 1.0, b->d, -1.0, w->Ap, w->r);
    if (replay) {
      cg_first_seq_kernel<<<1, kBlock, 0, s>>>(w->st, n, w->r);
    } else {
      dot_partial_kernel<<<vg, kBlock, 0, s>>>(n, w->r, w->r, w->partials);
      cg_first_fast_kernel<<<1, kBlock, 0, s>>>(w->st, w->partials, (int)vg);
    }
    MFE_LAUNCHED(ctx, "cg prologue");
    // Iteration 1: p = 1*r + 0*r.
    hipError_t he = enqueue_iteration(A, w, x->d, w->r, replay, s);
    if (he != hipSuccess) return hip_fail(ctx, he, "cg iteration 1");

    if (max_iter > 1) {
      if (eager) {
        for (int k = 2; k <= max_iter; ++k) {
          he = enqueue_iteration(A, w, x->d, w->p, replay, s);
          if (he != hipSuccess) return hip_fail(ctx, he, "cg iteration");
          if (k % kGraphChunk == 0) {
            MFE_CHECK(ctx, hipMemcpyAsync(w->h_st, w->st, sizeof(CgState), hipMemcpyDeviceToHost, s));
            MFE_CHECK(ctx, hipStreamSynchronize(s));
            if (w->h_st->done) break;
          }
        }
      } else {
        const bool reuse = w->exec && w->kA == A && w->kb == b && w->kx == x &&
                           w->khist == w->hist && w->kflags == (flags & MFE_CG_REPLAY) &&
                           w->kchunk == kGraphChunk && w->knnz == A->nnz;
        if (!reuse) {
          ws_drop_graph(w);
          MFE_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          hipError_t ce = hipSuccess;
          for (int c = 0; c < kGraphChunk && ce == hipSuccess; ++c)
            ce = enqueue_iteration(A, w, x->d, w->p, replay, s);
          hipGraph_t g = nullptr;
          hipError_t ee = hipStreamEndCapture(s, &g);
          if (ce != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return hip_fail(ctx, ce, "cg capture");
          }
          if (ee != hipSuccess) return hip_fail(ctx, ee, "hipStreamEndCapture");
          w->graph = g;
          MFE_CHECK(ctx, hipGraphInstantiate(&w->exec, g, nullptr, nullptr, 0));
          w->kA = A;
          w->kb = b;
          w->kx = x;
          w->khist = w->hist;
          w->kflags = flags & MFE_CG_REPLAY;
          w->kchunk = kGraphChunk;
          w->knnz = A->nnz;
        }
        const int chunks = (int)cdiv(max_iter - 1, kGraphChunk);
        for (int c = 0; c < chunks; ++c) {
          MFE_CHECK(ctx, hipGraphLaunch(w->exec, s));
          MFE_CHECK(ctx, hipMemcpyAsync(w->h_st, w->st, sizeof(CgState), hipMemcpyDeviceToHost, s));
          if (c + 1 < chunks) {
            MFE_CHECK(ctx, hipStreamSynchronize(s));
            if (w->h_st->done) break;
          }
        }
      }
    }
    cg_finish_kernel<<<vg, kBlock, 0, s>>>(n, w->st, w->p, x->d);
    MFE_LAUNCHED(ctx, "cg_finish_kernel");
  }
  MFE_CHECK(ctx, hipMemcpyAsync(w->h_st, w->st, sizeof(CgState), hipMemcpyDeviceToHost, s));
  MFE_CHECK(ctx, hipEventRecord(ctx->t_stop, s));
  MFE_CHECK(ctx, hipStreamSynchronize(s));
  const CgState hs = *w->h_st;
  *num_iters = hs.num_iters;
  *normr = hs.normr;
  if (times) {
    float ms = 0;
    MFE_CHECK(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
    for (int i = 0; i < 5; ++i) times[i] = 0;
    times[MFE_T_TOTAL] = 1e-3 * ms;
  }
  const int rows = hs.num_iters + (hs.status == MFE_BREAKDOWN ? 1 : 0);
  if (history && rows > 0) {
    MFE_CHECK(ctx, hipMemcpy(history, w->hist, sizeof(double) * 4 * (size_t)rows,
                             hipMemcpyDeviceToHost));
  }
  return hs.status == MFE_BREAKDOWN ? MFE_BREAKDOWN : MFE_OK;
}

}  // extern "C"
