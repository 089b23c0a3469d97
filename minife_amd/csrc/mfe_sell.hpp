// SELL-64 slice layout with slot pairs, shared by the SpMV kernels and the
// batched solver.
//
// A slice holds 64 consecutive rows (one per lane of a wave) and has width w =
// its longest row.  Its w*64 entries start at o*64.  Slots are stored in pairs
// (2p, 2p+1) as [pair][lane][2], so one 16-byte value load and one 8-byte
// column load fetch two slots of a lane; an odd w keeps its last slot as
// [lane].  Entries past a row's length are padding (value 0, a valid column)
// and are skipped in the sum, which adds the row's entries in stored order --
// the reference matvec's arithmetic.  Arrays carry a 64-entry tail (value 0,
// column 0): the pair loads of a width-1 slice reach up to 128 entries past
// its start.
#pragma once

#include "mfe_common.hpp"

namespace mfe {

constexpr int kSellTail = 64;

__host__ __device__ inline int64_t sell_pos(int64_t o64, int j, int w, int lane) {
  const int np = w >> 1;
  return j < 2 * np ? o64 + (int64_t)(j >> 1) * 128 + lane * 2 + (j & 1)
                    : o64 + (int64_t)np * 128 + lane;
}

template <typename T>
struct SellVec2;
template <>
struct SellVec2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct SellVec2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
typedef int sell_int2 __attribute__((ext_vector_type(2)));

// The sum over the lane's row (len entries) of vals * x[cols], added onto
// `sum` in stored order.  All value/column loads are issued first (pair index
// clamped into the slice), then all x gathers, then the additions.
template <typename T>
__device__ __forceinline__ T sell_row_sum(const int* cols, const T* vals, const T* x, int64_t o64,
                                          int w, int lane, int len, T sum) {
  using V2 = typename SellVec2<T>::type;
  if (w <= 0) return sum;
  const int np = w >> 1;
  const int last = np > 0 ? np - 1 : 0;
  int c[kSellMaxWidth];
  T v[kSellMaxWidth], xv[kSellMaxWidth];
#pragma unroll
  for (int p = 0; p < kSellMaxWidth / 2; ++p) {
    const int64_t at = o64 + (int64_t)(p < last ? p : last) * 128 + lane * 2;
    const sell_int2 cc = *reinterpret_cast<const sell_int2*>(cols + at);
    const V2 vv = *reinterpret_cast<const V2*>(vals + at);
    c[2 * p] = cc.x;
    c[2 * p + 1] = cc.y;
    v[2 * p] = vv.x;
    v[2 * p + 1] = vv.y;
  }
  {
    const int64_t at = (w & 1) ? o64 + (int64_t)np * 128 + lane : o64 + lane * 2;
    c[kSellMaxWidth - 1] = cols[at];
    v[kSellMaxWidth - 1] = vals[at];
  }
#pragma unroll
  for (int j = 0; j < kSellMaxWidth; ++j) xv[j] = x[c[j]];
  const int js = (w & 1) ? 2 * np : -1;  // the slot held in the last register
#pragma unroll
  for (int j = 0; j < kSellMaxWidth; ++j)
    if (j < len) {
      const T a = j == js ? v[kSellMaxWidth - 1] : v[j];
      const T b = j == js ? xv[kSellMaxWidth - 1] : xv[j];
      sum += a * b;
    }
  return sum;
}

}  // namespace mfe
