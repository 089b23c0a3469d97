// SELL-64 slice layout with slot pairs, shared by the SpMV kernels and the
// batched solver.
//
// A slice holds 64 consecutive rows (one per lane of a wave) and has width w =
// its longest row.  Its w*64 entries start at o*64.  Slots are stored in pairs
// as [pair][lane][2], so one 16-byte value load and one 8-byte column load
// fetch two slots of a lane.  An even w is w/2 pairs (slots 2p, 2p+1); an odd
// w keeps slot 0 alone as [lane] and pairs slots (2p+1, 2p+2) after it.  So
// in either case register j of the unrolled row holds slot j, and the width's
// parity picks one of two fully unrolled load paths (uniform per wave).
// Entries past a row's length are padding (value 0, a valid column) and are
// skipped in the sum, which adds the row's entries in stored order -- the
// reference matvec's arithmetic.  Arrays carry a tail of kSellTail entries
// (value 0, column 0): the clamped pair loads of a width-1 slice reach up to
// 192 entries past its start.
#pragma once

#include "mfe_common.hpp"

namespace mfe {

constexpr int kSellTail = 192;

__host__ __device__ inline int64_t sell_pos(int64_t o64, int j, int w, int lane) {
  if (w & 1) {
    if (j == 0) return o64 + lane;
    const int jj = j - 1;
    return o64 + 64 + (int64_t)(jj >> 1) * 128 + lane * 2 + (jj & 1);
  }
  return o64 + (int64_t)(j >> 1) * 128 + lane * 2 + (j & 1);
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
template <typename T, bool ODD>
__device__ __forceinline__ T sell_row_sum_w(const int* cols, const T* vals, const T* x, int64_t o64,
                                            int w, int lane, int len, T sum) {
  using V2 = typename SellVec2<T>::type;
  constexpr int kPairs = (kSellMaxWidth - (ODD ? 1 : 0)) / 2;  // 13 either way for 27
  constexpr int kFirst = ODD ? 1 : 0;                          // register of pair 0's first slot
  const int np = ODD ? (w - 1) >> 1 : w >> 1;
  const int last = np > 0 ? np - 1 : 0;
  const int64_t pb = o64 + (ODD ? 64 : 0);
  int c[kSellMaxWidth];
  T v[kSellMaxWidth], xv[kSellMaxWidth];
  if (ODD) {
    c[0] = cols[o64 + lane];
    v[0] = vals[o64 + lane];
  }
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    const int64_t at = pb + (int64_t)(p < last ? p : last) * 128 + lane * 2;
    const sell_int2 cc = *reinterpret_cast<const sell_int2*>(cols + at);
    const V2 vv = *reinterpret_cast<const V2*>(vals + at);
    c[kFirst + 2 * p] = cc.x;
    c[kFirst + 2 * p + 1] = cc.y;
    v[kFirst + 2 * p] = vv.x;
    v[kFirst + 2 * p + 1] = vv.y;
  }
  constexpr int kUsed = kFirst + 2 * kPairs;  // 27 (odd) or 26 (even)
#pragma unroll
  for (int j = 0; j < kUsed; ++j) xv[j] = x[c[j]];
#pragma unroll
  for (int j = 0; j < kUsed; ++j)
    if (j < len) sum += v[j] * xv[j];
  return sum;
}

template <typename T>
__device__ __forceinline__ T sell_row_sum(const int* cols, const T* vals, const T* x, int64_t o64,
                                          int w, int lane, int len, T sum) {
  if (w <= 0) return sum;
  return (w & 1) ? sell_row_sum_w<T, true>(cols, vals, x, o64, w, lane, len, sum)
                 : sell_row_sum_w<T, false>(cols, vals, x, o64, w, lane, len, sum);
}

typedef unsigned short sell_us2 __attribute__((ext_vector_type(2)));

// A lane's slots with 16-bit columns, loaded by the same instructions for
// either parity of the width: s = the single slot 0 of an odd width (for an
// even width a load of pair 0's address, unused), p[k] = pair k (index
// clamped into the slice).  Kept raw so that no wait for the loads is needed
// before the values are used (sell_slot picks slot j afterwards).
template <typename T>
struct SellLocalRaw {
  int odd;
  unsigned short c0;
  T v0;
  sell_us2 cp[kSellMaxWidth / 2];
  typename SellVec2<T>::type vp[kSellMaxWidth / 2];
};

// Values of the slice at o64, local columns of the slice at lo64 of lcols (the
// same position, or a zero region for slices whose columns are not read).
template <typename T>
__device__ __forceinline__ void sell_load_local(const unsigned short* lcols, int64_t lo64,
                                                const T* vals, int64_t o64, int w, int lane,
                                                SellLocalRaw<T>& r) {
  using V2 = typename SellVec2<T>::type;
  constexpr int kPairs = kSellMaxWidth / 2;  // 13: slots 1..26 (odd) or 0..25 (even)
  r.odd = w & 1;
  const int np = (w - r.odd) >> 1;
  const int last = np > 0 ? np - 1 : 0;
  const int first = r.odd ? 64 : 0;
  r.c0 = lcols[lo64 + lane];
  r.v0 = vals[o64 + lane];
#pragma unroll
  for (int p = 0; p < kPairs; ++p) {
    const int64_t d = first + (int64_t)(p < last ? p : last) * 128 + lane * 2;
    r.cp[p] = *reinterpret_cast<const sell_us2*>(lcols + lo64 + d);
    r.vp[p] = *reinterpret_cast<const V2*>(vals + o64 + d);
  }
}

// Slot j (a compile-time index after unrolling) of a raw load: column and
// value.  Odd widths: slot 0 single, slots 2k+1, 2k+2 = pair k; even widths:
// slots 2k, 2k+1 = pair k (slot 26 does not exist: column 0, value 0).
template <typename T>
__device__ __forceinline__ void sell_slot(const SellLocalRaw<T>& r, int j, int& c, T& v) {
  constexpr int kPairs = kSellMaxWidth / 2;
  int ce, co;
  T ve, vo;
  if (j == 0) {
    co = r.c0;
    vo = r.v0;
    ce = r.cp[0].x;
    ve = r.vp[0].x;
  } else if (j & 1) {
    const int k = (j - 1) >> 1;
    co = r.cp[k].x;
    vo = r.vp[k].x;
    ce = r.cp[k].y;
    ve = r.vp[k].y;
  } else {
    const int k = (j - 2) >> 1;
    co = r.cp[k].y;
    vo = r.vp[k].y;
    ce = (j >> 1) < kPairs ? (int)r.cp[j >> 1].x : 0;
    ve = (j >> 1) < kPairs ? r.vp[j >> 1].x : T(0);
  }
  c = r.odd ? co : ce;
  v = r.odd ? vo : ve;
}

}  // namespace mfe
