// Device side of the linked-cell search: cell lookup and the reference-order
// neighbour visit, shared by the pair export and the resident stepper.
#pragma once
#include <math.h>

#include "tsph_common.hpp"
#include "tsph_internal.hpp"

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
TSPH_D T dist2_pt(const T* xi, const T* xj, const T* box) {
#pragma clang fp contract(off)
  T acc = T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    T df = xi[d] - xj[d];
    if (PER) df = df - box[d] * rint_t<T>(df / box[d]);
    const T sq = df * df;
    acc = d == 0 ? sq : acc + sq;
  }
  return acc;
}
template <typename T, int D, bool PER>
TSPH_D T dist2(const T* __restrict__ pos, int64_t i, int64_t j, const T* box) {
  return dist2_pt<T, D, PER>(pos + i * D, pos + j * D, box);
}

TSPH_D int64_t find_cell(const int64_t* __restrict__ ukeys, int64_t m, int64_t key) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ukeys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < m && ukeys[lo] == key) ? lo : -1;
}

// Morton (Z-order) key of non-negative cell coordinates: the bits of the D
// axes interleaved, axis 0 the most significant of each group.  Consecutive
// keys form compact blocks of cells, so rows sorted by it keep every cell's
// 3^D neighbourhood close in memory (up to 21 bits per axis in 3-D, 31 in 2-D).
TSPH_HD uint64_t spread_bits2(uint64_t v) {
  v &= 0x7fffffffull;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}
TSPH_HD uint64_t spread_bits3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x001f00000000ffffull;
  v = (v | (v << 16)) & 0x001f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
template <int D>
TSPH_HD int64_t morton_key(const int64_t* c) {
  if constexpr (D == 1) return c[0];
  else if constexpr (D == 2) return (int64_t)((spread_bits2((uint64_t)c[0]) << 1) | spread_bits2((uint64_t)c[1]));
  else
    return (int64_t)((spread_bits3((uint64_t)c[0]) << 2) | (spread_bits3((uint64_t)c[1]) << 1) |
                     spread_bits3((uint64_t)c[2]));
}

struct GridDesc {
  int64_t ext[3];      // open: cells per axis after normalisation; periodic: ncell
  int64_t stride[3];
  int morton = 0;      // keys are morton_key of the cell (else row-major)
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
    int64_t key = 0, nbc[3] = {0, 0, 0};
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
      nbc[d] = nb < 0 ? 0 : nb;
      key += nbc[d] * g.stride[d];
    }
    if (!inside) continue;
    if (g.morton) key = morton_key<D>(nbc);
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

// ---------------------------------------------------------------- neighbour sources
// The pair sweeps take their candidates from one of two sources with the same
// visiting order, so a sweep gives the same bits over either:
//   NbList  per-target CSR lists (built by for_neighbours at the list radius);
//   NbCells a walk over the 3^D cells around the target's cell (edge >= list
//           radius) through a dense table of per-cell row ranges.
// Both visit the neighbour cells with offsets in lexicographic order (last
// axis fastest) and, in each cell, its fluid rows then its wall rows in
// ascending order, skipping the target itself.  The cell walk also offers the
// candidates beyond the list radius; every sweep re-tests the kernel support,
// which is inside that radius, so those contribute nothing.  Lanes of a wave
// that share a cell read the same rows, so the walk's loads are mostly
// broadcasts where a list walk gathers one cache line per lane.
// largest dense cell table the steppers allocate (16 B per cell)
constexpr int64_t kMaxDenseCells = 1ll << 27;

struct NbList {
  const int64_t* off;
  const int32_t* idx;
  // with_walls is implied by how the lists were built
  template <int D, typename F>
  TSPH_D void visit(int64_t i, bool /*with_walls*/, F&& f) const {
    const int64_t e1 = off[i + 1];
    for (int64_t e = off[i]; e < e1; ++e) f((int64_t)idx[e]);
  }
  // f(j, true) for every entry, U at a time in one straight-line block
  template <int D, int U = 4, typename F>
  TSPH_D void visit_masked(int64_t i, bool /*with_walls*/, F&& f) const {
    const int64_t e1 = off[i + 1];
    int64_t e = off[i];
    for (; e + U <= e1; e += U) {
      int32_t j[U];
#pragma unroll
      for (int u = 0; u < U; ++u) j[u] = idx[e + u];
#pragma unroll
      for (int u = 0; u < U; ++u) f((int64_t)j[u], true);
    }
    for (; e < e1; ++e) f((int64_t)idx[e], true);
  }
};

struct NbCells {
  const int4* R;            // per row-major cell: fluid rows [x, y), wall rows [z, w)
  const int64_t* cells;     // (rows, D) cell of every row at the last binning
  int64_t ext[3], stride[3];
  int periodic;
  // f runs for every candidate of the walk and re-tests the support itself.
  // A two-phase variant (scan, keep accepted rows, then run f over them) was
  // measured 2.6-3.7x slower on MI355X: it turns the shared candidate reads
  // of lanes in one cell into per-lane gathers again.
  // g(begin, end) for every non-empty candidate row range, in the order of visit
  template <int D, typename G>
  TSPH_D void ranges(int64_t i, bool with_walls, G&& g) const {
    int64_t ci[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) ci[d] = cells[i * D + d];
    constexpr int NOFF = D == 1 ? 3 : (D == 2 ? 9 : 27);
    for (int o = 0; o < NOFF; ++o) {
      int rem = o;
      int64_t L = 0;
      bool inside = true;
#pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        int64_t nb = ci[d] + (rem % 3 - 1);
        rem /= 3;
        if (periodic) {
          if (nb < 0) nb += ext[d];
          else if (nb >= ext[d]) nb -= ext[d];
        } else if (nb < 0 || nb >= ext[d]) {
          inside = false;
        }
        L += nb * stride[d];
      }
      if (!inside) continue;
      const int4 r = R[L];
      if (r.y > r.x) g(r.x, r.y);
      if (with_walls && r.w > r.z) g(r.z, r.w);
    }
  }
  template <int D, typename F>
  TSPH_D void visit(int64_t i, bool with_walls, F&& f) const {
    ranges<D>(i, with_walls, [&](int32_t b, int32_t e) {
      for (int32_t j = b; j < e; ++j)
        if (j != i) f((int64_t)j);
    });
  }
};

// Per-row lists in fixed slots interleaved by wave, four entries per lane: the
// K slots of a row (K a multiple of 4) form K / 4 chunks of four int32, and
// chunk c of row r sits at int4 index (r / 64) * 16 * K + 64 c + r % 64.  The
// 64 rows of a wave read their c-th chunks (entries 4c .. 4c + 3) as one
// coalesced 1 KB run, one 16-byte load per lane.  Filled in the cell walk's
// visiting order with the build radius test (k_slot_lists).  Fluid targets
// list fluid and wall rows, wall targets fluid rows only; with_walls == false
// skips the wall rows (index >= nf) of a fluid target's list.
struct NbSlots {
  const int32_t* idx;
  const int32_t* cnt;
  int64_t K, nf;
  TSPH_D const int4* chunks(int64_t i) const {
    return reinterpret_cast<const int4*>(idx) + (i >> 6) * 16 * K + (i & 63);
  }
  template <int D, typename F>
  TSPH_D void visit(int64_t i, bool with_walls, F&& f) const {
    const int4* p = chunks(i);
    const int n = cnt[i];
    for (int k = 0; k < n; k += 4) {
      const int4 q = p[(k >> 2) * 64];
      const int32_t j4[4] = {q.x, q.y, q.z, q.w};
      for (int u = 0; u < 4 && k + u < n; ++u) {
        const int64_t j = j4[u];
        if (!with_walls && j >= nf) continue;
        f(j);
      }
    }
  }
  // f(j4, keep4) once per chunk of four entries (CH chunks per block): the
  // chunk loads come first, then f evaluates the four entries together.
  // Slots past the row's count, and the wall rows with_walls == false skips,
  // hold the row itself with keep false.
  template <int D, int CH = 1, typename F>
  TSPH_D void visit_chunks(int64_t i, bool with_walls, F&& f) const {
    const int4* p = chunks(i);
    const int n = cnt[i];
    const int4 self = make_int4((int)i, (int)i, (int)i, (int)i);
    for (int k = 0; k < n; k += 4 * CH) {
      int4 q[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) q[c] = k + 4 * c < n ? p[((k >> 2) + c) * 64] : self;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int32_t raw[4] = {q[c].x, q[c].y, q[c].z, q[c].w};
        int32_t j4[4];
        bool keep4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool in = k + 4 * c + u < n;
          j4[u] = in ? raw[u] : (int32_t)i;
          keep4[u] = in && (with_walls || j4[u] < nf);
        }
        f(j4, keep4);
      }
    }
  }
  // f(j, keep) for every entry, one chunk of four at a time in one
  // straight-line block: the chunk load and the gathers f makes for its four
  // entries are all issued before the arithmetic that needs them.  Slots past
  // the row's count run f on the row itself with keep false, as do the wall
  // rows that with_walls == false skips; f must not branch on keep.
  // CH chunks (4 CH entries) per block.
  template <int D, int CH = 1, typename F>
  TSPH_D void visit_masked(int64_t i, bool with_walls, F&& f) const {
    const int4* p = chunks(i);
    const int n = cnt[i];
    const int4 self = make_int4((int)i, (int)i, (int)i, (int)i);
    for (int k = 0; k < n; k += 4 * CH) {
      int4 q[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) q[c] = k + 4 * c < n ? p[((k >> 2) + c) * 64] : self;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int32_t j4[4] = {q[c].x, q[c].y, q[c].z, q[c].w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool in = k + 4 * c + u < n;
          const int64_t j = in ? (int64_t)j4[u] : i;
          f(j, in && (with_walls || j < nf));
        }
      }
    }
  }
};

// Appends a row's entries to its slots in the NbSlots layout, four at a time:
// entries gather in registers and each full chunk is one 16-byte store; the
// last partial chunk is filled up with the row itself.  push_if keeps the
// select chain branch-free; only the store of a completed chunk branches.
struct SlotWriter {
  int4* p;          // chunk 0 of the row
  int4 buf;
  int k;
  int32_t self;
  TSPH_D SlotWriter(int32_t* idx, int64_t K, int64_t i)
      : p(reinterpret_cast<int4*>(idx) + (i >> 6) * 16 * K + (i & 63)),
        buf(make_int4((int)i, (int)i, (int)i, (int)i)), k(0), self((int32_t)i) {}
  TSPH_D void push_if(bool keep, int32_t j) {
    const int sl = k & 3;
    buf.x = keep && sl == 0 ? j : buf.x;
    buf.y = keep && sl == 1 ? j : buf.y;
    buf.z = keep && sl == 2 ? j : buf.z;
    buf.w = keep && sl == 3 ? j : buf.w;
    if (keep && sl == 3) p[(k >> 2) * 64] = buf;
    k += keep ? 1 : 0;
  }
  TSPH_D int finish() {
    if (k & 3) {
      const int sl = k & 3;
      buf.y = sl > 1 ? buf.y : self;
      buf.z = sl > 2 ? buf.z : self;
      buf.w = self;
      p[(k >> 2) * 64] = buf;
    }
    return k;
  }
};

// Dense per-cell row ranges of rows sorted by cell, fluid rows [0, nf) before
// wall rows [nf, nt): every run of equal cells writes its begin / end.  The
// table starts zeroed (k_cell_table_clear), so empty cells are empty ranges.
static __global__ void k_cell_table_clear(int4* __restrict__ R, int64_t ncells) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) R[c] = make_int4(0, 0, 0, 0);
}

template <int D>
__global__ void k_cell_table(const int64_t* __restrict__ cells, int64_t nt, int64_t nf, int64_t s0, int64_t s1,
                             int64_t s2, int4* __restrict__ R) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nt) return;
  const int64_t st[3] = {s0, s1, s2};
  auto lin = [&](int64_t k) {
    int64_t L = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) L += cells[k * D + d] * st[d];
    return L;
  };
  const int64_t L = lin(r);
  const bool wall = r >= nf;
  const bool first = r == 0 || r == nf || lin(r - 1) != L;
  const bool last = r == nt - 1 || r + 1 == nf || lin(r + 1) != L;
  int* cell = reinterpret_cast<int*>(R + L);
  if (first) cell[wall ? 2 : 0] = (int)r;
  if (last) cell[wall ? 3 : 1] = (int)(r + 1);
}

// FP32 copies of the build-time positions relative to row 0, for the list
// build's prefilter: float2 rows in 2-D, float4 (w unused) in 3-D.
template <int D> struct PosF;
template <> struct PosF<2> {
  using T = float2;
  static TSPH_D float at(const T& v, int d) { return d == 0 ? v.x : v.y; }
};
template <> struct PosF<3> {
  using T = float4;
  static TSPH_D float at(const T& v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
};

template <int D>
__global__ void k_pos_f32(const double* __restrict__ pos, int64_t nt, typename PosF<D>::T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  float v[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < D; ++d) v[d] = (float)(pos[i * D + d] - pos[d]);
  if constexpr (D == 2) out[i] = make_float2(v[0], v[1]);
  else out[i] = make_float4(v[0], v[1], v[2], 0.f);
}

// Per-row lists at the build radius (cell walk, the unfused distance test of
// the pair search, minimum image when PER) of rows [i0, nt).  Every pass writes each row's
// entry count, their maximum (kmax) and, if total is given, their sum; with
// FILL the entries go into K slots per row (NbSlots layout; a row with more
// than K entries keeps its first K and its full count, so the caller widens
// the slots and runs the pass again).  Each lane gathers four entries in
// registers and stores them as one 16-byte chunk; the slots of a row's last
// chunk past its count hold the row itself.  Candidates are fetched four at a
// time ahead of their tests; launched with grid_xcd.
//
// With PF the candidates are first tested in FP32 on posf: a squared distance
// below lo2 accepts, above hi2 rejects, and only the band between them, which
// holds the FP32 rounding error of the test, is decided by the FP64 test.
template <int D, bool FILL, bool PER, bool PF>
__global__ void __launch_bounds__(kBlock) k_slot_lists(NbCells walk, const double* __restrict__ pos,
                                                       const typename PosF<D>::T* __restrict__ posf, int64_t nf,
                                                       int64_t i0, int64_t nt, double rc2, float lo2, float hi2,
                                                       double b0, double b1, double b2, int32_t* __restrict__ cnt,
                                                       int32_t* __restrict__ idx, int64_t K,
                                                       unsigned int* __restrict__ kmax,
                                                       unsigned long long* __restrict__ total) {
  using VF = typename PosF<D>::T;
  const int64_t i = i0 + xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  int k = 0;
  if (i < nt) {
    const double box[3] = {b0, b1, b2};
    const float boxf[3] = {(float)b0, (float)b1, (float)b2};
    const float ibf[3] = {PER ? 1.0f / (float)b0 : 0.f, PER ? 1.0f / (float)b1 : 0.f,
                          PER && D > 2 ? 1.0f / (float)b2 : 0.f};
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = pos[i * D + d];
    VF xfi{};
    if constexpr (PF) xfi = posf[i];
    int4* p = FILL ? reinterpret_cast<int4*>(idx) + (i >> 6) * 16 * K + (i & 63) : nullptr;
    const int chunks = (int)(K >> 2);
    int4 buf = make_int4((int)i, (int)i, (int)i, (int)i);
    auto accept = [&](int32_t j) {
      if (FILL) {
        const int s = k & 3;
        buf.x = s == 0 ? j : buf.x;
        buf.y = s == 1 ? j : buf.y;
        buf.z = s == 2 ? j : buf.z;
        buf.w = s == 3 ? j : buf.w;
        if (s == 3 && (k >> 2) < chunks) p[(k >> 2) * 64] = buf;
      }
      ++k;
    };
    walk.ranges<D>(i, i < nf, [&](int32_t b, int32_t e) {
      for (int32_t j0 = b; j0 < e; j0 += 4) {
        if constexpr (PF) {
          VF c[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = posf[j0 + u < e ? j0 + u : e - 1];   // in range: e > j0 >= b
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int32_t j = j0 + u;
            if (j >= e || j == i) continue;
            float d2 = 0.f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              float df = PosF<D>::at(xfi, d) - PosF<D>::at(c[u], d);
              if (PER) df = df - boxf[d] * rintf(df * ibf[d]);
              d2 += df * df;
            }
            bool in = d2 < lo2;
            if (!in && d2 <= hi2) {
              double xj[D];
#pragma unroll
              for (int d = 0; d < D; ++d) xj[d] = pos[(int64_t)j * D + d];
              in = dist2_pt<double, D, PER>(xi, xj, box) <= rc2;
            }
            if (in) accept(j);
          }
        } else {
          double xj[4][D];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int64_t j = j0 + u < e ? j0 + u : e - 1;   // in range: e > j0 >= b
#pragma unroll
            for (int d = 0; d < D; ++d) xj[u][d] = pos[j * D + d];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int32_t j = j0 + u;
            if (j >= e || j == i) continue;
            if (dist2_pt<double, D, PER>(xi, xj[u], box) > rc2) continue;
            accept(j);
          }
        }
      }
    });
    if (FILL && (k & 3) && (k >> 2) < chunks) {
      const int s = k & 3;                      // the row itself in the unused slots
      buf.y = s > 1 ? buf.y : (int)i;
      buf.z = s > 2 ? buf.z : (int)i;
      buf.w = (int)i;
      p[(k >> 2) * 64] = buf;
    }
    cnt[i] = k;
  }
  int m = k;
  unsigned long long sum = (unsigned long long)k;
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(m, o, 64);
    m = v > m ? v : m;
    sum += __shfl_xor(sum, o, 64);
  }
  // one atomic per block: per-wave atomics on one word serialise
  __shared__ int bm[kBlock / 64];
  __shared__ unsigned long long bs[kBlock / 64];
  if ((threadIdx.x & 63) == 0) { bm[threadIdx.x >> 6] = m; bs[threadIdx.x >> 6] = sum; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { m = bm[w] > m ? bm[w] : m; sum += bs[w]; }
    if ((unsigned int)m > __atomic_load_n(kmax, __ATOMIC_RELAXED)) atomicMax(kmax, (unsigned int)m);
    if (total && sum) atomicAdd(total, sum);
  }
}

// Builds the NbSlots lists of rows [0, nt) over walk: one pass at the current
// width K (a counting pass while K is 0), and when a list did not fit, the
// slots are widened to its length + 1/8 + 4 (a multiple of 4) and the pass
// runs again.  total, if given, receives the entry count.  Reads the longest
// list back, so the stream is synchronised.
//
// posf (scratch) enables the FP32 prefilter.  extent bounds |x - x_0| over the
// rows in every axis (x_0 = row 0).  Each FP32 coordinate then carries at most
// e0 = 2^-23 extent of rounding, a candidate difference at most
// e = 2 e0 + 2^-24 1.01 rc (+ 2^-22 L for the periodic image), and the squared
// distance at most 2 sqrt(D) rc e + D e^2 + 4.4 2^-24 rc^2 near rc^2.  Twice
// that is the undecided band; wider than 1 % of rc^2, the build runs in FP64.
template <int D, bool PER>
int build_slot_lists(const NbCells& walk, const double* pos, int64_t nf, int64_t nt, double rc2,
                     const double* box, DevBuf& cnt, DevBuf& idx, DevBuf& kmax, int64_t& K,
                     unsigned long long* total, hipStream_t stream, DevBuf* posf = nullptr,
                     double extent = 0.0) {
  TSPH_TRY(cnt.ensure(sizeof(int32_t) * (size_t)(nt > 0 ? nt : 1)));
  TSPH_TRY(kmax.ensure(sizeof(unsigned int)));
  const unsigned g = grid_xcd(nt);
  const double b0 = PER ? box[0] : 0.0, b1 = PER ? box[1] : 0.0, b2 = PER && D > 2 ? box[2] : 0.0;
  const double rc = sqrt(rc2);
  double e = 2.0 * ldexp(extent, -23) + ldexp(1.01 * rc, -24);
  if (PER) e += ldexp(fmax(fmax(b0, b1), b2), -22);
  const double band = 2.0 * (2.0 * sqrt((double)D) * rc * e + D * e * e + 4.4 * ldexp(rc2, -24));
  const bool pf = posf && extent > 0.0 && band < 0.01 * rc2;
  float lo2 = 0.f, hi2 = 0.f;
  using VF = typename PosF<D>::T;
  if (pf) {
    lo2 = nextafterf((float)(rc2 - band), 0.0f);
    hi2 = nextafterf((float)(rc2 + band), INFINITY);
    TSPH_TRY(posf->ensure(sizeof(VF) * (size_t)(nt > 0 ? nt : 1)));
    k_pos_f32<D><<<grid_for(nt), kBlock, 0, stream>>>(pos, nt, posf->as<VF>());
    TSPH_CHECK_LAUNCH();
  }
  const VF* pf_ptr = pf ? posf->as<VF>() : nullptr;
  for (int pass = 0; pass < 3; ++pass) {
    TSPH_CHECK_HIP(hipMemsetAsync(kmax.ptr, 0, sizeof(unsigned int), stream));
    if (total) TSPH_CHECK_HIP(hipMemsetAsync(total, 0, sizeof(unsigned long long), stream));
    int32_t* out = nullptr;
    if (K > 0) {
      TSPH_TRY(idx.ensure(sizeof(int32_t) * (size_t)((nt + 63) / 64) * 64 * (size_t)K));
      out = idx.as<int32_t>();
    }
#define TSPH_SLOT_PASS(FILL, PFV)                                                                     \
  k_slot_lists<D, FILL, PER, PFV><<<g, kBlock, 0, stream>>>(walk, pos, pf_ptr, nf, 0, nt, rc2, lo2, hi2, b0, b1, b2, \
                                                            cnt.as<int32_t>(), out, K, kmax.as<unsigned int>(), total)
    if (K > 0 && pf) TSPH_SLOT_PASS(true, true);
    else if (K > 0) TSPH_SLOT_PASS(true, false);
    else if (pf) TSPH_SLOT_PASS(false, true);
    else TSPH_SLOT_PASS(false, false);
#undef TSPH_SLOT_PASS
    TSPH_CHECK_LAUNCH();
    unsigned int longest = 0;
    TSPH_CHECK_HIP(hipMemcpyAsync(&longest, kmax.ptr, sizeof(longest), hipMemcpyDeviceToHost, stream));
    TSPH_CHECK_HIP(hipStreamSynchronize(stream));
    if (K > 0 && (int64_t)longest <= K) return TSPH_OK;
    K = ((int64_t)longest + longest / 8 + 4 + 3) / 4 * 4;
  }
  TSPH_REQUIRE(false, TSPH_ERR_HIP, "neighbour lists did not fit %lld slots per row", (long long)K);
  return TSPH_OK;
}

}  // namespace tsph
