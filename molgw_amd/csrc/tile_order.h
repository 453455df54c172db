// Device-side block ordering shared by the MFMA kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mgw {

constexpr int NXCD = 8;   // MI355X: workgroup w is dispatched to XCD w % 8

// Renumber so that consecutive logical ids run on the same XCD and share
// its L2: bijective on [0, G).
__device__ inline int64_t xcd_logical(int64_t w, int64_t G) {
  const int64_t q = G / NXCD, r = G % NXCD;
  const int64_t x = w % NXCD, i = w / NXCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Lower-triangle tile t (0-based, of T x T tiles) -> (bm, bn), bm >= bn, in a
// grouped order: super-rows of TRI_GROUP tile rows, walked column by column,
// so that concurrent workgroups share TRI_GROUP row panels and few column
// panels (a row-by-row walk would give every tile a fresh column panel).
constexpr int TRI_GROUP = 8;
__device__ inline void tri_grouped(int t, int T, int& bm, int& bn) {
  int r0 = 0;
  for (;;) {
    const int r1 = min(r0 + TRI_GROUP, T);
    if (r1 * (r1 + 1) / 2 > t || r1 == T) break;
    r0 = r1;
  }
  const int g = min(TRI_GROUP, T - r0);
  int u = t - r0 * (r0 + 1) / 2;
  if (u < r0 * g) {
    bn = u / g;
    bm = r0 + u % g;
    return;
  }
  u -= r0 * g;
  int c = 0;
  while (c < g - 1 && u >= g - c) {
    u -= g - c;
    ++c;
  }
  bn = r0 + c;
  bm = r0 + c + u;
}

}  // namespace mgw
