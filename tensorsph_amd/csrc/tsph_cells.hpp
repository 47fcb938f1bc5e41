// Device side of the linked-cell search: cell lookup and the reference-order
// neighbour visit, shared by the pair export and the resident stepper.
#pragma once
#include "tsph_common.hpp"

namespace tsph {

template <typename T> TSPH_D T floor_t(T x);
template <> TSPH_D double floor_t<double>(double x) { return floor(x); }
template <> TSPH_D float floor_t<float>(float x) { return floorf(x); }
template <typename T> TSPH_D T rint_t(T x);
template <> TSPH_D double rint_t<double>(double x) { return rint(x); }
template <> TSPH_D float rint_t<float>(float x) { return rintf(x); }

// ---------------------------------------------------------------- distances
// Unfused, in T, summed in axis order: the reference's
// (pos[i] - pos[j]).square().sum(dim=1).
template <typename T, int D, bool PER>
TSPH_D T dist2(const T* __restrict__ pos, int64_t i, int64_t j, const T* box) {
#pragma clang fp contract(off)
  T acc = T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    T df = pos[i * D + d] - pos[j * D + d];
    if (PER) df = df - box[d] * rint_t<T>(df / box[d]);
    const T sq = df * df;
    acc = d == 0 ? sq : acc + sq;
  }
  return acc;
}

TSPH_D int64_t find_cell(const int64_t* __restrict__ ukeys, int64_t m, int64_t key) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ukeys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < m && ukeys[lo] == key) ? lo : -1;
}

struct GridDesc {
  int64_t ext[3];      // open: cells per axis after normalisation; periodic: ncell
  int64_t stride[3];
};

// Visits every j in the 3^D neighbour cells of particle i (reference order:
// offsets lexicographic, members in stable cell order) that passes the
// distance test, calling f(j).
template <typename T, int D, bool PER, typename F>
TSPH_D void for_neighbours(const T* __restrict__ pos, int64_t i, const int64_t* __restrict__ cells,
                           const GridDesc& g, const int64_t* __restrict__ ukeys,
                           const int64_t* __restrict__ ustart, const int64_t* __restrict__ order,
                           int64_t ncells, T c2, const T* box, F&& f) {
  int64_t ci[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) ci[d] = cells[i * D + d];
  constexpr int NOFF = D == 1 ? 3 : (D == 2 ? 9 : 27);
  for (int o = 0; o < NOFF; ++o) {
    int rem = o;
    int64_t off[3] = {0, 0, 0};
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { off[d] = rem % 3 - 1; rem /= 3; }
    int64_t key = 0;
    bool inside = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int64_t nb = ci[d] + off[d];
      if (PER) {
        nb %= g.ext[d];
        if (nb < 0) nb += g.ext[d];
      } else if (nb < 0 || nb >= g.ext[d]) {
        inside = false;
      }
      key += (nb < 0 ? 0 : nb) * g.stride[d];
    }
    if (!inside) continue;
    const int64_t slot = find_cell(ukeys, ncells, key);
    if (slot < 0) continue;
    const int64_t s0 = ustart[slot], s1 = ustart[slot + 1];
    for (int64_t s = s0; s < s1; ++s) {
      const int64_t j = order[s];
      if (j == i) continue;
      if (dist2<T, D, PER>(pos, i, j, box) <= c2) f(j);
    }
  }
}

}  // namespace tsph
