// Binning of the resident stepper's rows into a fixed cell grid by a stable
// counting sort.
//
// The grid (cells of edge rc, the list radius) covers the clamp box and every
// row's position at the last full build, so it stays valid while the fluid
// moves inside the box: a list rebuild re-bins only the fluid rows and the
// walls keep their rows.  Cells are ordered by a blocked Morton key (blocks
// of 16 x 16 cells in 2-D, 8 x 8 x 8 in 3-D, row-major over blocks, Morton
// inside), which keeps consecutive rows spatially compact over a dense key
// range of about the grid's size.  The sort: one pass counts the rows of
// every key, an exclusive scan gives each key its first row, a scatter puts
// the rows of a key into its range in arrival order, and a rank pass orders
// each range by the rows' previous index.  The result is the stable sort of
// the rows by key, the same on every run.  A row outside the grid is binned
// into the nearest cell and raises a flag, after which the caller regrids.
#pragma once
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "tsph_cells.hpp"
#include "tsph_common.hpp"

namespace tsph {

// What a list build reports (written by the device into pinned host memory
// at the end of the build): the longest list, the largest halo, the tiles
// with a halo, whether a row was binned from outside the grid, and the slot
// width K the build ran with.
struct BuildStats {
  unsigned int kmax, hmax, used, outside, K;
};

template <int D> struct BinOrder {
  static constexpr int LB = D == 2 ? 4 : 3;   // log2 of the block edge in cells
  static constexpr int B = 1 << LB;
  static constexpr int64_t PER_BLOCK = D == 2 ? B * B : B * B * B;
};

struct BinGrid {
  double rc;               // cell edge
  int64_t cmin[3];         // floor(x / rc) of the grid's first cell
  int64_t ext[3];          // cells per axis
  int64_t nb[3];           // key blocks per axis
  int64_t M;               // dense key range
};

TSPH_HD uint64_t compact_bits2(uint64_t v) {
  v &= 0x5555555555555555ull;
  v = (v | (v >> 1)) & 0x3333333333333333ull;
  v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
  v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
  v = (v | (v >> 16)) & 0x00000000ffffffffull;
  return v;
}

template <int D> TSPH_D int64_t bin_key(const int64_t* c, const BinGrid& g) {
  using O = BinOrder<D>;
  int64_t blk = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) blk = blk * g.nb[d] + (c[d] >> O::LB);
  int64_t lo[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) lo[d] = c[d] & (O::B - 1);
  return blk * O::PER_BLOCK + morton_key<D>(lo);
}

// the cell of a key (inverse of bin_key)
template <int D> TSPH_D void bin_cell(int64_t key, const BinGrid& g, int64_t* c) {
  using O = BinOrder<D>;
  int64_t blk = key / O::PER_BLOCK;
  const uint64_t m = (uint64_t)(key % O::PER_BLOCK);
  int64_t lo[3] = {0, 0, 0};
  if constexpr (D == 2) {
    lo[0] = (int64_t)compact_bits2(m >> 1);
    lo[1] = (int64_t)compact_bits2(m);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      uint64_t v = 0;
#pragma unroll
      for (int b = 0; b < O::LB; ++b) v |= ((m >> (3 * b + (2 - d))) & 1ull) << b;
      lo[d] = (int64_t)v;
    }
  }
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    c[d] = (blk % g.nb[d]) * O::B + lo[d];
    blk /= g.nb[d];
  }
}

// Runs of equal keys among the live lanes of a wave (lanes [0, live)): head
// is the first lane of this lane's run, len the run's length.
TSPH_D void key_run(int32_t k, int64_t live, int& head, int& len) {
  const int lane = threadIdx.x & 63;
  const int32_t prev = __shfl_up(k, 1, 64);
  const bool is_head = lane == 0 || prev != k;
  const unsigned long long heads = __ballot(is_head);
  const unsigned long long below = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);
  head = 63 - __clzll(below);                          // last head at or below this lane
  const unsigned long long above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
  int end = above ? __ffsll(above) - 1 : 64;           // next head above this lane
  const int nlive = live < 64 ? (int)live : 64;
  end = end < nlive ? end : nlive;
  len = end - head;
}

// Cell (grid-relative, clamped into the grid) and key of rows [r0, r1); the
// rows of every key are counted.  *outside is raised for a row outside.
template <int D>
__global__ void k_bin(const double* __restrict__ pos, int64_t r0, int64_t r1, BinGrid g,
                      int64_t* __restrict__ cells, int32_t* __restrict__ keys, int32_t* __restrict__ count,
                      unsigned int* __restrict__ outside) {
  const int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= r1) return;
  int64_t c[3] = {0, 0, 0};
  bool out = false;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int64_t v = (int64_t)floor(pos[r * D + d] / g.rc) - g.cmin[d];
    if (v < 0) { v = 0; out = true; }
    if (v >= g.ext[d]) { v = g.ext[d] - 1; out = true; }
    c[d] = v;
    cells[r * D + d] = v;
  }
  const int32_t k = (int32_t)bin_key<D>(c, g);
  keys[r] = k;
  // rows of one cell mostly arrive together: one atomic per run of equal
  // keys in the wave
  int head, len;
  key_run(k, r1 - (r - (int64_t)(threadIdx.x & 63)), head, len);
  if ((int)(threadIdx.x & 63) == head) atomicAdd(&count[k], len);
  if (out) *outside = 1u;
}

// Exclusive scan of count[0, m) into start[0, m] (start[m] = total) in three
// passes of kScanTile entries per block.
constexpr int kScanPer = 16;
constexpr int kScanTile = kBlock * kScanPer;

TSPH_D int scan_block_excl(int v, int* wtot, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  int base = 0, all = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
    base += k < w ? wtot[k] : 0;
    all += wtot[k];
  }
  *total = all;
  __syncthreads();
  return base + inc - v;
}

static __global__ void __launch_bounds__(kBlock) k_scan_sums(const int32_t* __restrict__ count, int64_t m,
                                                      int32_t* __restrict__ part) {
  __shared__ int wtot[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) s += base + q < m ? count[base + q] : 0;
  int total;
  (void)scan_block_excl(s, wtot, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// one block: part[0, nparts) to its exclusive prefix
static __global__ void __launch_bounds__(kBlock) k_scan_parts(int32_t* __restrict__ part, int64_t nparts) {
  __shared__ int wtot[kBlock / 64];
  int carry = 0;
  for (int64_t b0 = 0; b0 < nparts; b0 += kBlock) {
    const int64_t b = b0 + threadIdx.x;
    const int v = b < nparts ? part[b] : 0;
    int total;
    const int ex = scan_block_excl(v, wtot, &total);
    if (b < nparts) part[b] = carry + ex;
    carry += total;
  }
}

static __global__ void __launch_bounds__(kBlock) k_scan_apply(const int32_t* __restrict__ count, int64_t m,
                                                       const int32_t* __restrict__ part, int32_t base_row,
                                                       int32_t* __restrict__ start) {
  __shared__ int wtot[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  int loc[kScanPer], s = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    loc[q] = base + q < m ? count[base + q] : 0;
    s += loc[q];
  }
  int total;
  int run = scan_block_excl(s, wtot, &total) + part[blockIdx.x] + base_row;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    if (base + q < m) start[base + q] = run;
    run += loc[q];
    if (base + q == m - 1) start[m] = run;
  }
}

// rows [r0, r1) into their key ranges, in arrival order (count is consumed)
static __global__ void k_bin_scatter(const int32_t* __restrict__ keys, int64_t r0, int64_t r1,
                              const int32_t* __restrict__ start, int32_t* __restrict__ count,
                              int32_t* __restrict__ tmp) {
  const int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= r1) return;
  const int32_t k = keys[r];
  int head, len;
  key_run(k, r1 - (r - (int64_t)(threadIdx.x & 63)), head, len);
  const int lane = threadIdx.x & 63;
  int32_t old = 0;
  if (lane == head) old = atomicSub(&count[k], len);
  old = __shfl(old, head, 64);
  tmp[start[k] + old - len + (lane - head)] = (int32_t)r;
}

// each key range ordered by previous row: new2old[slot] for slots [s0, s1)
static __global__ void k_bin_rank(const int32_t* __restrict__ tmp, const int32_t* __restrict__ keys, int64_t s0,
                           int64_t s1, const int32_t* __restrict__ start, int64_t* __restrict__ new2old) {
  const int64_t s = s0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= s1) return;
  const int32_t r = tmp[s];
  const int32_t k = keys[r];
  const int32_t a = start[k], b = start[k + 1];
  int rank = 0;
  for (int32_t t = a; t < b; ++t) rank += tmp[t] < r ? 1 : 0;
  new2old[a + rank] = r;
}

// floor(x / rc) bounds of rows [0, nt): per block the smallest and largest
// cell per axis into blk[6 * block + d] / blk[6 * block + 3 + d] (axes past D:
// 0), then k_bin_bounds_reduce into out[6].
template <int D>
__global__ void __launch_bounds__(kBlock) k_bin_bounds(const double* __restrict__ pos, int64_t nt, double rc,
                                                       long long* __restrict__ blk) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  long long lo[3] = {LLONG_MAX, LLONG_MAX, LLONG_MAX}, hi[3] = {LLONG_MIN, LLONG_MIN, LLONG_MIN};
  if (r < nt) {
#pragma unroll
    for (int d = 0; d < D; ++d) lo[d] = hi[d] = (long long)floor(pos[r * D + d] / rc);
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    for (int o = 32; o > 0; o >>= 1) {
      const long long a = __shfl_xor(lo[d], o, 64), b = __shfl_xor(hi[d], o, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  __shared__ long long part[6][kBlock / 64];
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) { part[d][threadIdx.x >> 6] = lo[d]; part[3 + d][threadIdx.x >> 6] = hi[d]; }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    long long v = part[q][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = q < 3 ? (part[q][w] < v ? part[q][w] : v)
                                                               : (part[q][w] > v ? part[q][w] : v);
    blk[6 * (int64_t)blockIdx.x + q] = (q % 3) < D ? v : 0;
  }
}

static __global__ void __launch_bounds__(kBlock) k_bin_bounds_reduce(const long long* __restrict__ blk, int64_t nblk,
                                                              long long* __restrict__ out) {
  long long v[6] = {LLONG_MAX, LLONG_MAX, LLONG_MAX, LLONG_MIN, LLONG_MIN, LLONG_MIN};
  for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const long long x = blk[6 * b + q];
      v[q] = q < 3 ? (x < v[q] ? x : v[q]) : (x > v[q] ? x : v[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    for (int o = 32; o > 0; o >>= 1) {
      const long long x = __shfl_xor(v[q], o, 64);
      v[q] = q < 3 ? (x < v[q] ? x : v[q]) : (x > v[q] ? x : v[q]);
    }
  }
  __shared__ long long part[6][kBlock / 64];
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) part[q][threadIdx.x >> 6] = v[q];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    long long r = part[q][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      r = q < 3 ? (part[q][w] < r ? part[q][w] : r) : (part[q][w] > r ? part[q][w] : r);
    out[q] = r;
  }
}

}  // namespace tsph
