// Many right-hand sides on one matrix: Y = A X for X with K columns (row-major
// n x K, K = 4, 8 or 16), and K simultaneous FP64 CG solves that share one
// SpMM per iteration (each column keeps its own alpha, beta and stopping test,
// cg_solve.hpp:69-229 per column, with tree-reduced dots).
//
// SpMM tile: L = K / CPL lanes per row, each lane owns CPL consecutive columns
// of the row (one 8*CPL-byte load per non-zero), so a row's K columns are one
// 8K-byte stretch of X per non-zero.  The L lanes of a row read the same
// value/column entry (a broadcast).  The row is walked in chunks of U entries
// whose loads are all issued before the first product; each column's sum runs
// in stored column order, so column k of Y equals A * X[:, k] from the
// single-vector SpMV bit for bit.  The autotuner times every (CPL, U) tile on
// the caller's matrix and keeps the fastest for that matrix and K.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mfe_common.hpp"

namespace mfe {

struct SpmmTile {
  int cpl, unroll;
};

// Candidate tiles per K (index = tile id).
static const SpmmTile kTiles16[] = {{1, 9}, {2, 9}, {4, 9}, {8, 9},
                                    {2, 3}, {4, 3}, {8, 3}, {16, 3}};
static const SpmmTile kTiles8[] = {{1, 9}, {2, 9}, {4, 9}, {8, 9}, {2, 3}, {4, 3}, {8, 3}};
static const SpmmTile kTiles4[] = {{1, 9}, {2, 9}, {4, 9}, {2, 3}, {4, 3}};

static int tile_count(int K) {
  return K == 16 ? (int)(sizeof(kTiles16) / sizeof(SpmmTile))
                 : K == 8 ? (int)(sizeof(kTiles8) / sizeof(SpmmTile))
                          : (int)(sizeof(kTiles4) / sizeof(SpmmTile));
}
static SpmmTile tile_of(int K, int id) {
  return K == 16 ? kTiles16[id] : K == 8 ? kTiles8[id] : kTiles4[id];
}

// Sum over the lanes of one wave that hold the same column (lane % L), in a
// fixed butterfly order; valid in every lane.
template <int L>
__device__ __forceinline__ double sum_same_column(double v) {
#pragma unroll
  for (int off = kWave / 2; off >= L; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Y = A X; with partials non-null also the workgroup partials of Y[:,k].P[:,k]
// (partials[block * K + k]), P = X here (the CG's p.Ap).
template <int K, int CPL, int U>
__global__ __launch_bounds__(kBlock) void spmm_kernel(int n, const int* __restrict__ ro,
                                                      const int* __restrict__ cols,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ X,
                                                      double* __restrict__ Y,
                                                      double* __restrict__ partials,
                                                      const int* __restrict__ done) {
  if (done && *done) return;
  constexpr int L = K / CPL;
  constexpr int RPB = kBlock / L;  // rows per workgroup
  const int t = (int)threadIdx.x;
  const int row = (int)blockIdx.x * RPB + t / L;
  const int k0 = (t % L) * CPL;
  double acc[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) acc[q] = 0.0;
  if (row < n) {
    const int a = ro[row], len = ro[row + 1] - a;
    for (int j0 = 0; j0 < len; j0 += U) {
      int c[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = a + min(j0 + u, len - 1);
        c[u] = cols[j];
        v[u] = vals[j];
      }
      double xv[U][CPL];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double* src = X + (int64_t)c[u] * K + k0;
        if constexpr (CPL % 2 == 0) {
#pragma unroll
          for (int q = 0; q < CPL; q += 2) {
            const double2 d = *reinterpret_cast<const double2*>(src + q);
            xv[u][q] = d.x;
            xv[u][q + 1] = d.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < CPL; ++q) xv[u][q] = src[q];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j0 + u < len) {
#pragma unroll
          for (int q = 0; q < CPL; ++q) acc[q] += v[u] * xv[u][q];
        }
    }
    double* dst = Y + (int64_t)row * K + k0;
#pragma unroll
    for (int q = 0; q < CPL; ++q) dst[q] = acc[q];
  }
  if (partials) {
    __shared__ double red[kBlock / kWave][K];
    double term[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q)
      term[q] = row < n ? acc[q] * X[(int64_t)row * K + k0 + q] : 0.0;
    const int lane = t % kWave, wave = t / kWave;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const double s = sum_same_column<L>(term[q]);
      if (lane < L) red[wave][lane * CPL + q] = s;
    }
    __syncthreads();
    if (t < K) {
      double s = red[0][t];
#pragma unroll
      for (int w = 1; w < kBlock / kWave; ++w) s += red[w][t];
      partials[(int64_t)blockIdx.x * K + t] = s;
    }
  }
}

template <int K>
static hipError_t launch_spmm(int tile, int n, const int* ro, const int* cols, const double* vals,
                              const double* X, double* Y, double* partials, const int* done,
                              hipStream_t s, int* blocks_out) {
  const SpmmTile T = tile_of(K, tile);
  const int rpb = kBlock / (K / T.cpl);
  const unsigned g = (unsigned)cdiv(n, rpb);
  if (blocks_out) *blocks_out = (int)g;
  if (n == 0) return hipSuccess;
#define MFE_SPMM_CASE(C, UU)                                                                \
  if (T.cpl == C && T.unroll == UU) {                                                      \
    spmm_kernel<K, (C <= K ? C : K), UU><<<g, kBlock, 0, s>>>(n, ro, cols, vals, X, Y,      \
                                                               partials, done);            \
    return hipGetLastError();                                                              \
  }
  MFE_SPMM_CASE(1, 9)
  MFE_SPMM_CASE(2, 9)
  MFE_SPMM_CASE(4, 9)
  MFE_SPMM_CASE(8, 9)
  MFE_SPMM_CASE(2, 3)
  MFE_SPMM_CASE(4, 3)
  MFE_SPMM_CASE(8, 3)
  MFE_SPMM_CASE(16, 3)
#undef MFE_SPMM_CASE
  return hipErrorInvalidValue;
}

static hipError_t launch_spmm_k(int K, int tile, int n, const int* ro, const int* cols,
                                const double* vals, const double* X, double* Y, double* partials,
                                const int* done, hipStream_t s, int* blocks_out) {
  switch (K) {
    case 16: return launch_spmm<16>(tile, n, ro, cols, vals, X, Y, partials, done, s, blocks_out);
    case 8: return launch_spmm<8>(tile, n, ro, cols, vals, X, Y, partials, done, s, blocks_out);
    case 4: return launch_spmm<4>(tile, n, ro, cols, vals, X, Y, partials, done, s, blocks_out);
    default: return hipErrorInvalidValue;
  }
}

static int k_slot(int K) { return K == 16 ? 2 : K == 8 ? 1 : 0; }

// ---------------------------------------------------------- autotuner -------

struct SpmmTuning {
  const mfe_csr* A = nullptr;
  uint64_t gen = 0;
  int tile[3] = {-1, -1, -1};
  double ms[3][16] = {};
};

// Times every tile for K on A (3 launches each after one warm-up, HIP events
// on ctx's stream) and keeps the fastest.  X and Y are scratch of n*K.
static int tune_spmm(mfe_ctx* ctx, const mfe_csr* A, int K, const double* X, double* Y,
                     SpmmTuning& tun) {
  hipStream_t s = ctx->stream;
  const int nt = tile_count(K);
  const int slot = k_slot(K);
  int best = 0;
  double best_ms = 1e300;
  for (int id = 0; id < nt; ++id) {
    MFE_CHECK(ctx, launch_spmm_k(K, id, (int)A->nrows, A->row_offsets, A->cols, A->vals, X, Y,
                                 nullptr, nullptr, s, nullptr));
    MFE_CHECK(ctx, hipEventRecord(ctx->t_start, s));
    const int reps = 3;
    for (int r = 0; r < reps; ++r)
      MFE_CHECK(ctx, launch_spmm_k(K, id, (int)A->nrows, A->row_offsets, A->cols, A->vals, X, Y,
                                   nullptr, nullptr, s, nullptr));
    MFE_CHECK(ctx, hipEventRecord(ctx->t_stop, s));
    MFE_CHECK(ctx, hipEventSynchronize(ctx->t_stop));
    float ms = 0;
    MFE_CHECK(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
    tun.ms[slot][id] = ms / reps;
    if (tun.ms[slot][id] < best_ms) {
      best_ms = tun.ms[slot][id];
      best = id;
    }
  }
  tun.tile[slot] = best;
  return MFE_OK;
}

// ------------------------------------------------------ column reductions ---

constexpr int kColStage = 1024;  // partial rows per stage-1 workgroup

// out[b*K + k] = sum of in[r*K + k] over rows r in [b*chunk, (b+1)*chunk),
// fixed order: thread (t / K) walks rows t/K, t/K + 256/K, ...; then the
// 256/K threads of a column add up in a fixed tree.
template <int K>
__global__ __launch_bounds__(kBlock) void reduce_cols_kernel(const double* __restrict__ in,
                                                             int rows, int chunk,
                                                             double* __restrict__ out,
                                                             const int* __restrict__ done) {
  if (done && *done) return;
  constexpr int G = kBlock / K;  // threads per column
  __shared__ double sh[kBlock];
  const int t = (int)threadIdx.x;
  const int k = t % K, g = t / K;
  const int r0 = (int)blockIdx.x * chunk, r1 = min(r0 + chunk, rows);
  double s = 0.0;
  for (int r = r0 + g; r < r1; r += G) s += in[(int64_t)r * K + k];
  sh[t] = s;
  __syncthreads();
  for (int w = G / 2; w > 0; w >>= 1) {
    if (g < w) sh[t] += sh[t + w * K];
    __syncthreads();
  }
  if (g == 0) out[(int64_t)blockIdx.x * K + k] = sh[k];
}

// ------------------------------------------------------------ CG state -------

struct MultiState {
  double rtrans[16], oldrtrans[16], pAp[16], alpha[16], beta[16], normr[16], tol[16], bnorm[16];
  int active[16], iters[16], breakdown[16];
  int k, max_iter, done, pad;
};

__global__ void multi_init_kernel(MultiState* st, int K, int max_iter, double rel_tol,
                                  const double* bb, const double* rr) {
  const int c = (int)threadIdx.x;
  if (c == 0) {
    st->k = 1;
    st->max_iter = max_iter;
  }
  if (c < 16) {
    const bool live = c < K;
    st->rtrans[c] = live ? rr[c] : 0.0;
    st->oldrtrans[c] = 0.0;
    st->normr[c] = live ? sqrt(rr[c]) : 0.0;
    st->bnorm[c] = live ? sqrt(bb[c]) : 0.0;
    st->tol[c] = rel_tol * st->bnorm[c];
    st->alpha[c] = st->beta[c] = st->pAp[c] = 0.0;
    st->iters[c] = 0;
    st->breakdown[c] = 0;
    st->active[c] = live && max_iter >= 1 && st->normr[c] > st->tol[c];
  }
  __syncthreads();
  if (c == 0) {
    int any = 0;
    for (int q = 0; q < K; ++q) any |= st->active[q];
    st->done = !any;
  }
}

// P = R (k = 1) or P = R + beta P for active columns.
template <int K>
__global__ __launch_bounds__(kBlock) void multi_p_kernel(int64_t nk, const MultiState* st,
                                                         const double* __restrict__ R,
                                                         double* __restrict__ P) {
  if (st->done) return;
  const bool first = st->k == 1;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += stride) {
    const int c = (int)(i % K);
    if (!st->active[c]) continue;
    P[i] = first ? 1.0 * R[i] + 0.0 * R[i] : 1.0 * R[i] + st->beta[c] * P[i];
  }
}

__global__ void multi_alpha_kernel(MultiState* st, int K, const double* pAp) {
  if (st->done) return;
  const int c = (int)threadIdx.x;
  if (c < K && st->active[c]) {
    st->pAp[c] = pAp[c];
    if (!(pAp[c] > 0.0)) {  // non-positive curvature: stop this column
      st->breakdown[c] = 1;
      st->active[c] = 0;
      st->alpha[c] = 0.0;
    } else {
      st->alpha[c] = st->rtrans[c] / pAp[c];
    }
  }
}

// X += alpha P, R -= alpha AP for active columns, with R.R partials per column
// (grid-stride: a thread always sees the same column since the stride is a
// multiple of K).
template <int K>
__global__ __launch_bounds__(kBlock) void multi_update_kernel(int64_t nk, const MultiState* st,
                                                              double* __restrict__ X,
                                                              double* __restrict__ R,
                                                              const double* __restrict__ P,
                                                              const double* __restrict__ AP,
                                                              double* __restrict__ partials) {
  if (st->done) return;
  const int t = (int)threadIdx.x;
  const int c = t % K;
  const bool act = st->active[c];
  const double a = st->alpha[c];
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + t; i < nk; i += stride) {
    double r = R[i];
    if (act) {
      X[i] = 1.0 * X[i] + a * P[i];
      r = 1.0 * r + (-a) * AP[i];
      R[i] = r;
    }
    s += r * r;
  }
  __shared__ double red[kBlock / kWave][K];
  s = sum_same_column<K>(s);
  const int lane = t % kWave, wave = t / kWave;
  if (lane < K) red[wave][lane] = s;
  __syncthreads();
  if (t < K) {
    double v = red[0][t];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; ++w) v += red[w][t];
    partials[(int64_t)blockIdx.x * K + t] = v;
  }
}

__global__ void multi_next_kernel(MultiState* st, int K, const double* rr) {
  if (st->done) return;
  const int c = (int)threadIdx.x;
  const int k = st->k;
  if (c < K && st->active[c]) {
    st->oldrtrans[c] = st->rtrans[c];
    st->rtrans[c] = rr[c];
    st->normr[c] = sqrt(rr[c]);
    st->beta[c] = st->rtrans[c] / st->oldrtrans[c];
    st->iters[c] = k;
    if (!(st->normr[c] > st->tol[c]) || k + 1 > st->max_iter) st->active[c] = 0;
  }
  __syncthreads();
  if (c == 0) {
    int any = 0;
    for (int q = 0; q < K; ++q) any |= st->active[q];
    st->k = k + 1;
    st->done = !any;
  }
}

// R = B - Y elementwise.
__global__ __launch_bounds__(kBlock) void multi_resid_kernel(int64_t nk, const double* B,
                                                             const double* Y, double* R) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += stride)
    R[i] = 1.0 * B[i] + (-1.0) * Y[i];
}

// partials of column dots a[:,k].b[:,k]
template <int K>
__global__ __launch_bounds__(kBlock) void multi_dot_kernel(int64_t nk, const double* a,
                                                           const double* b, double* partials) {
  const int t = (int)threadIdx.x;
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + t; i < nk; i += stride) s += a[i] * b[i];
  __shared__ double red[kBlock / kWave][K];
  s = sum_same_column<K>(s);
  const int lane = t % kWave, wave = t / kWave;
  if (lane < K) red[wave][lane] = s;
  __syncthreads();
  if (t < K) {
    double v = red[0][t];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; ++w) v += red[w][t];
    partials[(int64_t)blockIdx.x * K + t] = v;
  }
}

struct MultiWorkspace {
  int64_t nk = -1;
  int K = 0;
  double *R = nullptr, *P = nullptr, *AP = nullptr, *part = nullptr, *part2 = nullptr,
         *res = nullptr;
  int64_t part_rows = 0;
  MultiState* st = nullptr;
  MultiState* h_st = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const void *kA = nullptr, *kX = nullptr;
  uint64_t kgen = 0;
  int ktile = -1;
  SpmmTuning tun;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    kA = kX = nullptr;
  }
  void release() {
    drop_graph();
    (void)hipFree(R);
    (void)hipFree(P);
    (void)hipFree(AP);
    (void)hipFree(part);
    (void)hipFree(part2);
    (void)hipFree(res);
    (void)hipFree(st);
    if (h_st) (void)hipHostFree(h_st);
    R = P = AP = part = part2 = res = nullptr;
    st = nullptr;
    h_st = nullptr;
    nk = -1;
  }
};

static unsigned multi_vec_grid(int64_t nk) {
  const int64_t b = cdiv(nk, kBlock);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, kVecBlocks));
}

static int multi_prepare(mfe_ctx* ctx, int64_t n, int K, MultiWorkspace** out) {
  if (!ctx->multi) ctx->multi = new MultiWorkspace();
  MultiWorkspace* w = ctx->multi;
  const int64_t nk = n * K;
  if (w->nk != nk || w->K != K) {
    w->release();
    const size_t bytes = sizeof(double) * (size_t)std::max<int64_t>(nk, 1);
    MFE_CHECK(ctx, hipMalloc(&w->R, bytes));
    MFE_CHECK(ctx, hipMalloc(&w->P, bytes));
    MFE_CHECK(ctx, hipMalloc(&w->AP, bytes));
    // partial rows: the SpMM's workgroups (at most n with one row each) or
    // the vector kernels' grid
    w->part_rows = std::max<int64_t>(n, kVecBlocks) + 1;
    MFE_CHECK(ctx, hipMalloc(&w->part, sizeof(double) * (size_t)(w->part_rows * K)));
    MFE_CHECK(ctx, hipMalloc(&w->part2, sizeof(double) * (size_t)((cdiv(w->part_rows, kColStage) + 1) * K)));
    MFE_CHECK(ctx, hipMalloc(&w->res, sizeof(double) * 64));
    MFE_CHECK(ctx, hipMalloc(&w->st, sizeof(MultiState)));
    MFE_CHECK(ctx, hipHostMalloc(&w->h_st, sizeof(MultiState), hipHostMallocDefault));
    w->nk = nk;
    w->K = K;
  }
  *out = w;
  return MFE_OK;
}

// Two-stage column reduction of part[rows][K] into out[K].
static hipError_t reduce_cols(int K, const double* part, int rows, double* part2, double* out,
                              const int* done, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, cdiv(rows, kColStage));
  switch (K) {
    case 16:
      reduce_cols_kernel<16><<<nb, kBlock, 0, s>>>(part, rows, kColStage, part2, done);
      reduce_cols_kernel<16><<<1, kBlock, 0, s>>>(part2, nb, nb, out, done);
      break;
    case 8:
      reduce_cols_kernel<8><<<nb, kBlock, 0, s>>>(part, rows, kColStage, part2, done);
      reduce_cols_kernel<8><<<1, kBlock, 0, s>>>(part2, nb, nb, out, done);
      break;
    default:
      reduce_cols_kernel<4><<<nb, kBlock, 0, s>>>(part, rows, kColStage, part2, done);
      reduce_cols_kernel<4><<<1, kBlock, 0, s>>>(part2, nb, nb, out, done);
  }
  return hipGetLastError();
}

template <int K>
static hipError_t enqueue_multi_iteration(MultiWorkspace* w, const mfe_csr* A, double* X,
                                          int tile, hipStream_t s) {
  const int64_t nk = w->nk;
  const unsigned vg = multi_vec_grid(nk);
  const int* done = &w->st->done;
  multi_p_kernel<K><<<vg, kBlock, 0, s>>>(nk, w->st, w->R, w->P);
  int nb = 0;
  hipError_t e = launch_spmm<K>(tile, (int)A->nrows, A->row_offsets, A->cols, A->vals, w->P,
                                w->AP, w->part, done, s, &nb);
  if (e != hipSuccess) return e;
  if ((e = reduce_cols(K, w->part, nb, w->part2, w->res, done, s)) != hipSuccess) return e;
  multi_alpha_kernel<<<1, 64, 0, s>>>(w->st, K, w->res);
  multi_update_kernel<K><<<vg, kBlock, 0, s>>>(nk, w->st, X, w->R, w->P, w->AP, w->part);
  if ((e = reduce_cols(K, w->part, (int)vg, w->part2, w->res + 16, done, s)) != hipSuccess)
    return e;
  multi_next_kernel<<<1, 64, 0, s>>>(w->st, K, w->res + 16);
  return hipGetLastError();
}

static hipError_t enqueue_multi(int K, MultiWorkspace* w, const mfe_csr* A, double* X, int tile,
                                hipStream_t s) {
  switch (K) {
    case 16: return enqueue_multi_iteration<16>(w, A, X, tile, s);
    case 8: return enqueue_multi_iteration<8>(w, A, X, tile, s);
    default: return enqueue_multi_iteration<4>(w, A, X, tile, s);
  }
}

static hipError_t launch_multi_dot(int K, int64_t nk, const double* a, const double* b,
                                   double* part, hipStream_t s) {
  const unsigned vg = multi_vec_grid(nk);
  switch (K) {
    case 16: multi_dot_kernel<16><<<vg, kBlock, 0, s>>>(nk, a, b, part); break;
    case 8: multi_dot_kernel<8><<<vg, kBlock, 0, s>>>(nk, a, b, part); break;
    default: multi_dot_kernel<4><<<vg, kBlock, 0, s>>>(nk, a, b, part);
  }
  return hipGetLastError();
}

static bool k_ok(int K) { return K == 4 || K == 8 || K == 16; }

static int ensure_tuned(mfe_ctx* ctx, const mfe_csr* A, int K, MultiWorkspace* w) {
  const int slot = k_slot(K);
  if (w->tun.A == A && w->tun.gen == A->plan_gen && w->tun.tile[slot] >= 0) return MFE_OK;
  if (w->tun.A != A || w->tun.gen != A->plan_gen) {
    w->tun = SpmmTuning();
    w->tun.A = A;
    w->tun.gen = A->plan_gen;
  }
  // scratch operands: R (filled with ones) and AP
  MFE_CHECK(ctx, hipMemsetAsync(w->R, 0, sizeof(double) * (size_t)w->nk, ctx->stream));
  return tune_spmm(ctx, A, K, w->R, w->AP, w->tun);
}

}  // namespace mfe

using namespace mfe;

extern "C" {

int mfe_spmm_tune(mfe_ctx* ctx, const mfe_csr* A, int nrhs, int* tile, double* ms, int* ntiles) {
  if (!ctx || !A || !k_ok(nrhs)) return MFE_BAD_ARG;
  MFE_CHECK(ctx, hipSetDevice(ctx->device));
  MultiWorkspace* w = nullptr;
  int st;
  if ((st = multi_prepare(ctx, A->nrows, nrhs, &w))) return st;
  w->tun = SpmmTuning();  // re-time on request
  if ((st = ensure_tuned(ctx, A, nrhs, w))) return st;
  const int slot = k_slot(nrhs);
  if (tile) *tile = w->tun.tile[slot];
  if (ntiles) *ntiles = tile_count(nrhs);
  if (ms)
    for (int i = 0; i < tile_count(nrhs); ++i) ms[i] = w->tun.ms[slot][i];
  return MFE_OK;
}

int mfe_spmm_tile_info(int nrhs, int tile, int* cols_per_lane, int* unroll, int* rows_per_block) {
  if (!k_ok(nrhs) || tile < 0 || tile >= tile_count(nrhs)) return MFE_BAD_ARG;
  const SpmmTile T = tile_of(nrhs, tile);
  if (cols_per_lane) *cols_per_lane = T.cpl;
  if (unroll) *unroll = T.unroll;
  if (rows_per_block) *rows_per_block = kBlock / (nrhs / T.cpl);
  return MFE_OK;
}

int mfe_spmm(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* X, mfe_vec* Y, int nrhs, int tile) {
  if (!ctx || !A || !X || !Y || !k_ok(nrhs)) return MFE_BAD_ARG;
  const int64_t nk = A->nrows * nrhs;
  if (X->n != nk || Y->n != nk) return fail(ctx, MFE_BAD_ARG, "spmm: X and Y need nrows*nrhs values");
  if (tile >= tile_count(nrhs)) return fail(ctx, MFE_BAD_ARG, "spmm: no such tile");
  MFE_CHECK(ctx, hipSetDevice(ctx->device));
  int st;
  if (tile < 0) {  // the tuned tile for this matrix and nrhs
    MultiWorkspace* w = nullptr;
    if ((st = multi_prepare(ctx, A->nrows, nrhs, &w)) || (st = ensure_tuned(ctx, A, nrhs, w)))
      return st;
    tile = w->tun.tile[k_slot(nrhs)];
  }
  MFE_CHECK(ctx, launch_spmm_k(nrhs, tile, (int)A->nrows, A->row_offsets, A->cols, A->vals, X->d,
                               Y->d, nullptr, nullptr, ctx->stream, nullptr));
  return MFE_OK;
}

int mfe_cg_solve_multi(mfe_ctx* ctx, const mfe_csr* A, const mfe_vec* B, mfe_vec* X, int nrhs,
                       int max_iter, double rel_tol, int* iters, double* rel_residual,
                       double* seconds) {
  if (!ctx || !A || !B || !X || !k_ok(nrhs) || max_iter < 0) return MFE_BAD_ARG;
  if (!A->has_local_indices)
    return fail(ctx, MFE_NOT_LOCAL,
                "cg_solve_multi: A.has_local_indices is false; call make_local_matrix(A) first");
  const int K = nrhs;
  const int64_t n = A->nrows, nk = n * K;
  if (B->n != nk || X->n != nk)
    return fail(ctx, MFE_BAD_ARG, "cg_solve_multi: B and X need nrows*nrhs values (row-major)");
  MFE_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  MultiWorkspace* w = nullptr;
  int st;
  if ((st = multi_prepare(ctx, n, K, &w)) || (st = ensure_tuned(ctx, A, K, w))) return st;
  const int tile = w->tun.tile[k_slot(K)];
  MFE_CHECK(ctx, hipEventRecord(ctx->t_start, s));
  if (n > 0) {
    // R = B - A X ; tolerances from ||B[:,k]||, rtrans = ||R[:,k]||^2
    int nb = 0;
    MFE_CHECK(ctx, launch_spmm_k(K, tile, (int)n, A->row_offsets, A->cols, A->vals, X->d, w->AP,
                                 nullptr, nullptr, s, &nb));
    multi_resid_kernel<<<multi_vec_grid(nk), kBlock, 0, s>>>(nk, B->d, w->AP, w->R);
    MFE_CHECK(ctx, launch_multi_dot(K, nk, B->d, B->d, w->part, s));
    MFE_CHECK(ctx, reduce_cols(K, w->part, (int)multi_vec_grid(nk), w->part2, w->res + 32, nullptr, s));
    MFE_CHECK(ctx, launch_multi_dot(K, nk, w->R, w->R, w->part, s));
    MFE_CHECK(ctx, reduce_cols(K, w->part, (int)multi_vec_grid(nk), w->part2, w->res + 48, nullptr, s));
    multi_init_kernel<<<1, 64, 0, s>>>(w->st, K, max_iter, rel_tol, w->res + 32, w->res + 48);
    MFE_LAUNCHED(ctx, "multi prologue");
    const bool reuse = w->exec && w->kA == A && w->kgen == A->plan_gen && w->kX == X->d &&
                       w->ktile == tile;
    if (!reuse) {
      w->drop_graph();
      MFE_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      for (int c = 0; c < 25 && ce == hipSuccess; ++c) ce = enqueue_multi(K, w, A, X->d, tile, s);
      hipGraph_t g = nullptr;
      hipError_t ee = hipStreamEndCapture(s, &g);
      if (ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(ctx, ce, "cg_solve_multi capture");
      }
      if (ee != hipSuccess) return hip_fail(ctx, ee, "hipStreamEndCapture");
      w->graph = g;
      MFE_CHECK(ctx, hipGraphInstantiate(&w->exec, g, nullptr, nullptr, 0));
      w->kA = A;
      w->kgen = A->plan_gen;
      w->kX = X->d;
      w->ktile = tile;
    }
    const int chunks = (int)cdiv(std::max(max_iter, 1), 25);
    for (int c = 0; c < chunks; ++c) {
      MFE_CHECK(ctx, hipGraphLaunch(w->exec, s));
      MFE_CHECK(ctx, hipMemcpyAsync(w->h_st, w->st, sizeof(MultiState), hipMemcpyDeviceToHost, s));
      MFE_CHECK(ctx, hipStreamSynchronize(s));
      if (w->h_st->done) break;
    }
  }
  MFE_CHECK(ctx, hipEventRecord(ctx->t_stop, s));
  MFE_CHECK(ctx, hipEventSynchronize(ctx->t_stop));
  float ms = 0;
  MFE_CHECK(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
  if (seconds) *seconds = 1e-3 * ms;
  if (n == 0) return MFE_OK;
  const MultiState hs = *w->h_st;
  int brk = 0;
  for (int c = 0; c < K; ++c) {
    if (iters) iters[c] = hs.iters[c];
    if (rel_residual) rel_residual[c] = hs.bnorm[c] > 0 ? hs.normr[c] / hs.bnorm[c] : hs.normr[c];
    brk |= hs.breakdown[c];
  }
  if (brk) return fail(ctx, MFE_BREAKDOWN, "cg_solve_multi: CG breakdown in a column");
  return MFE_OK;
}

void mfe_multi_workspace_free(mfe_ctx* ctx) {
  if (ctx && ctx->multi) {
    ctx->multi->release();
    delete ctx->multi;
    ctx->multi = nullptr;
  }
}

}  // extern "C"
