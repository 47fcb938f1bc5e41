// Linked-cell neighbour search on gfx950.
//
// The cell of a particle is floor(x / cutoff) - min (open domains,
// neighbor_search.py:65-68) or floor(x / width) mod ncell (periodic,
// threed/neighbors3.py:93-97), computed in the positions' own precision so the
// cell coordinates are the reference's bit for bit.  Particles are grouped by
// a stable radix sort of the row-major cell key (the reference's stable
// argsort), occupied cells are found by flagging key changes, and every
// particle scans the 3^D neighbour cells by binary search in the occupied
// keys (the reference's batched searchsorted).  Pair distances use the
// reference's unfused arithmetic: ((dx^2 + dy^2) + dz^2) <= (T)(cutoff^2).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdarg.h>
#include <string.h>

#include "tsph_common.hpp"
#include "tsph_internal.hpp"
#include "tsph_cells.hpp"

namespace tsph {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

constexpr unsigned long long kEnc = 1ull << 62;

// ---------------------------------------------------------------- cell coordinates
// Cell of every row, and per block the smallest and largest cell per axis as
// keys (c + kEnc) in blk[6 * block + d] / blk[6 * block + 3 + d] (~0 / 0 for
// a block without rows), stored plainly: thousands of blocks meeting on six
// words with atomics serialise.  k_minmax_reduce combines the records.
template <typename T, int D>
__global__ void k_cell_coords(const T* __restrict__ pos, int64_t n, T cutoff,
                              int64_t* __restrict__ cells, unsigned long long* __restrict__ blk) {
  __shared__ long long smin[D][kBlock / 64], smax[D][kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  long long lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { lo[d] = LLONG_MAX; hi[d] = LLONG_MIN; }
  if (i < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T len = pos[i * D + d] / cutoff;
      const long long c = (long long)floor_t<T>(len);
      cells[i * D + d] = c;
      lo[d] = c; hi[d] = c;
    }
  }
  // wave64 reduction, then across the block's waves
#pragma unroll
  for (int d = 0; d < D; ++d) {
    for (int off = 32; off > 0; off >>= 1) {
      long long a = __shfl_xor(lo[d], off, 64);
      long long b = __shfl_xor(hi[d], off, 64);
      lo[d] = a < lo[d] ? a : lo[d];
      hi[d] = b > hi[d] ? b : hi[d];
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) { smin[d][wave] = lo[d]; smax[d][wave] = hi[d]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* o = blk + 6 * (int64_t)blockIdx.x;
    for (int d = 0; d < 3; ++d) { o[d] = ~0ull; o[3 + d] = 0ull; }
    for (int d = 0; d < D; ++d) {
      long long a = LLONG_MAX, b = LLONG_MIN;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        a = smin[d][w] < a ? smin[d][w] : a;
        b = smax[d][w] > b ? smax[d][w] : b;
      }
      if (a <= b) {
        o[d] = (unsigned long long)(a + (long long)kEnc);
        o[3 + d] = (unsigned long long)(b + (long long)kEnc);
      }
    }
  }
}

// minmax[0..2] = smallest, minmax[3..5] = largest key over nblk records; one
// block of kBlock threads.
__global__ void __launch_bounds__(kBlock) k_minmax_reduce(const unsigned long long* __restrict__ blk,
                                                          int64_t nblk, unsigned long long* __restrict__ minmax) {
  unsigned long long v[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  for (int64_t b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const unsigned long long lo = blk[6 * b + q], hi = blk[6 * b + 3 + q];
      v[q] = lo < v[q] ? lo : v[q];
      v[3 + q] = hi > v[3 + q] ? hi : v[3 + q];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const unsigned long long lo = __shfl_xor(v[q], o, 64), hi = __shfl_xor(v[3 + q], o, 64);
      v[q] = lo < v[q] ? lo : v[q];
      v[3 + q] = hi > v[3 + q] ? hi : v[3 + q];
    }
  }
  __shared__ unsigned long long part[6][kBlock / 64];
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) part[q][threadIdx.x >> 6] = v[q];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    unsigned long long r = part[q][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      r = q < 3 ? (part[q][w] < r ? part[q][w] : r) : (part[q][w] > r ? part[q][w] : r);
    minmax[q] = r;
  }
}

template <typename T, int D>
__global__ void k_cell_coords_periodic(const T* __restrict__ pos, int64_t n, T w0, T w1, T w2,
                                       int64_t m0, int64_t m1, int64_t m2, int64_t* __restrict__ cells) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T width[3] = {w0, w1, w2};
  const int64_t ncell[3] = {m0, m1, m2};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T len = pos[i * D + d] / width[d];
    long long c = (long long)floor_t<T>(len);
    long long m = ncell[d];
    c %= m;
    if (c < 0) c += m;
    cells[i * D + d] = c;
  }
}

template <int D>
__global__ void k_keys(int64_t* __restrict__ cells, int64_t n, int64_t c0, int64_t c1, int64_t c2,
                       int64_t s0, int64_t s1, int64_t s2, int normalise, int morton,
                       int64_t* __restrict__ keys, int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t cm[3] = {c0, c1, c2}, st[3] = {s0, s1, s2};
  int64_t key = 0, cc[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int64_t c = cells[i * D + d];
    if (normalise) { c -= cm[d]; cells[i * D + d] = c; }
    cc[d] = c;
    key += c * st[d];
  }
  keys[i] = morton ? morton_key<D>(cc) : key;
  idx[i] = (int32_t)i;
}

__global__ void k_flag_starts(const int64_t* __restrict__ skeys, int64_t n, int32_t* __restrict__ flags,
                              const int32_t* __restrict__ order32, int64_t* __restrict__ order) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  flags[s] = (s == 0 || skeys[s] != skeys[s - 1]) ? 1 : 0;
  order[s] = order32[s];
}

__global__ void k_write_cells(const int64_t* __restrict__ skeys, const int32_t* __restrict__ flags,
                              const int32_t* __restrict__ cid, int64_t n,
                              int64_t* __restrict__ ukeys, int64_t* __restrict__ ustart) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (flags[s]) {
    const int64_t c = cid[s] - 1;
    ukeys[c] = skeys[s];
    ustart[c] = s;
  }
  if (s == n - 1) ustart[cid[s]] = n;
}

// MODE 0: count pairs j > i (kept by the fluid-count prune); MODE 1: write them.
template <typename T, int D, bool PER, int MODE>
__global__ void k_pairs(const T* __restrict__ pos, int64_t n, const int64_t* __restrict__ cells,
                        GridDesc g, const int64_t* __restrict__ ukeys,
                        const int64_t* __restrict__ ustart, const int64_t* __restrict__ order,
                        int64_t ncells, T c2, T b0, T b1, T b2, int64_t fluid_count,
                        int64_t* __restrict__ count, const int64_t* __restrict__ off,
                        int64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T box[3] = {b0, b1, b2};
  int64_t k = 0;
  int64_t base = MODE == 1 ? off[i] : 0;
  for_neighbours<T, D, PER>(pos, i, cells, g, ukeys, ustart, order, ncells, c2, box,
                            [&](int64_t j) {
                              if (j <= i) return;
                              if (fluid_count >= 0 && i >= fluid_count && j >= fluid_count) return;
                              if (MODE == 1) {
                                out[2 * (base + k)] = i;
                                out[2 * (base + k) + 1] = j;
                              }
                              ++k;
                            });
  if (MODE == 0) count[i] = k;
  if (MODE == 1) {
    // insertion sort of this particle's segment by j (lexicographic output)
    int64_t* seg = out + 2 * base;
    for (int64_t a = 1; a < k; ++a) {
      const int64_t v = seg[2 * a + 1];
      int64_t b = a - 1;
      while (b >= 0 && seg[2 * b + 1] > v) { seg[2 * (b + 1) + 1] = seg[2 * b + 1]; --b; }
      seg[2 * (b + 1) + 1] = v;
    }
  }
}

// ---------------------------------------------------------------- neighbour lists from pairs
__global__ void k_count_pairs(const int64_t* __restrict__ pairs, int64_t np, int64_t n_fluid,
                              int64_t n_total, int all_targets,
                              unsigned long long* __restrict__ cnt,
                              unsigned long long* __restrict__ wcnt, int* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
  if (i < 0 || j < 0 || i >= n_total || j >= n_total || i == j) { atomicOr(bad, 1); return; }
  if (all_targets || i < n_fluid) atomicAdd(&cnt[i], 1ull);
  if (all_targets || j < n_fluid) atomicAdd(&cnt[j], 1ull);
  if (!all_targets) {
    if (i >= n_fluid && j < n_fluid) atomicAdd(&wcnt[i - n_fluid], 1ull);
    if (j >= n_fluid && i < n_fluid) atomicAdd(&wcnt[j - n_fluid], 1ull);
  }
}

__global__ void k_fill_pairs(const int64_t* __restrict__ pairs, int64_t np, int64_t n_fluid,
                             int all_targets, const int64_t* __restrict__ off,
                             unsigned long long* __restrict__ cur, int32_t* __restrict__ idx,
                             const int64_t* __restrict__ woff,
                             unsigned long long* __restrict__ wcur, int32_t* __restrict__ widx) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
  if (all_targets || i < n_fluid) idx[off[i] + (int64_t)atomicAdd(&cur[i], 1ull)] = (int32_t)j;
  if (all_targets || j < n_fluid) idx[off[j] + (int64_t)atomicAdd(&cur[j], 1ull)] = (int32_t)i;
  if (!all_targets) {
    if (i >= n_fluid && j < n_fluid)
      widx[woff[i - n_fluid] + (int64_t)atomicAdd(&wcur[i - n_fluid], 1ull)] = (int32_t)j;
    if (j >= n_fluid && i < n_fluid)
      widx[woff[j - n_fluid] + (int64_t)atomicAdd(&wcur[j - n_fluid], 1ull)] = (int32_t)i;
  }
}

// Sorts each target's list ascending, so gathers run in a fixed order.
__global__ void k_sort_lists(const int64_t* __restrict__ off, int64_t nt, int32_t* __restrict__ idx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int32_t* seg = idx + off[t];
  const int64_t k = off[t + 1] - off[t];
  for (int64_t a = 1; a < k; ++a) {
    const int32_t v = seg[a];
    int64_t b = a - 1;
    while (b >= 0 && seg[b] > v) { seg[b + 1] = seg[b]; --b; }
    seg[b + 1] = v;
  }
}

// exclusive scan of counts (n) into off (n + 1); returns total on the host
int exclusive_scan(tsph_ctx* ctx, const unsigned long long* cnt, int64_t n, int64_t* off,
                   hipStream_t stream, int64_t* total) {
  TSPH_CHECK_HIP(hipMemsetAsync(off, 0, sizeof(int64_t), stream));
  if (n > 0) {
    size_t bytes = 0;
    TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt,
                                                    (unsigned long long*)(off + 1), (int)n, stream));
    TSPH_TRY(ctx->scan_tmp.ensure(bytes));
    TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(ctx->scan_tmp.ptr, bytes, cnt,
                                                    (unsigned long long*)(off + 1), (int)n, stream));
  }
  if (total) {
    TSPH_CHECK_HIP(hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  }
  return TSPH_OK;
}

// ---------------------------------------------------------------- search driver
template <typename T, int D>
int run_search(tsph_ctx* ctx, const T* pos, int64_t n, double cutoff, const double* box,
               hipStream_t stream, bool morton) {
  CellSearch& cs = ctx->search;
  cs.n = n; cs.dim = D; cs.periodic = box != nullptr; cs.ncells = 0; cs.morton = morton ? 1 : 0;
  if (n == 0) return TSPH_OK;
  TSPH_TRY(cs.cells.ensure(sizeof(int64_t) * n * D));
  TSPH_TRY(cs.keys.ensure(sizeof(int64_t) * n));
  TSPH_TRY(cs.skeys.ensure(sizeof(int64_t) * n));
  TSPH_TRY(cs.order.ensure(sizeof(int64_t) * n));
  TSPH_TRY(cs.order32.ensure(sizeof(int32_t) * n));
  TSPH_TRY(cs.idx_in.ensure(sizeof(int32_t) * n));
  TSPH_TRY(cs.flags.ensure(sizeof(int32_t) * 2 * n));
  TSPH_TRY(cs.ukeys.ensure(sizeof(int64_t) * n));
  TSPH_TRY(cs.ustart.ensure(sizeof(int64_t) * (n + 1)));
  TSPH_TRY(cs.minmax.ensure(sizeof(unsigned long long) * 6 + sizeof(T) * 3 + sizeof(int64_t) * 3));
  const unsigned grid = grid_for(n);
  if (!box) {
    TSPH_TRY(cs.mm_blk.ensure(sizeof(unsigned long long) * 6 * (size_t)grid));
    k_cell_coords<T, D><<<grid, kBlock, 0, stream>>>(pos, n, (T)cutoff, cs.cells.as<int64_t>(),
                                                     cs.mm_blk.as<unsigned long long>());
    k_minmax_reduce<<<1, kBlock, 0, stream>>>(cs.mm_blk.as<unsigned long long>(), (int64_t)grid,
                                              cs.minmax.as<unsigned long long>());
    TSPH_CHECK_LAUNCH();
    unsigned long long mm[6];
    TSPH_CHECK_HIP(hipMemcpyAsync(mm, cs.minmax.ptr, sizeof(mm), hipMemcpyDeviceToHost, stream));
    TSPH_CHECK_HIP(hipStreamSynchronize(stream));
    double cells_total = 1.0;
    for (int d = 0; d < 3; ++d) { cs.cmin[d] = 0; cs.ext[d] = 1; }
    for (int d = 0; d < D; ++d) {
      cs.cmin[d] = (int64_t)(mm[d] - kEnc);
      const int64_t cmax = (int64_t)(mm[3 + d] - kEnc);
      cs.ext[d] = cmax - cs.cmin[d] + 1;
      cells_total *= (double)cs.ext[d];
    }
    TSPH_REQUIRE(cells_total < 4.0e18, TSPH_ERR_ARG,
                 "cell grid of %.3g cells overflows the int64 key; the cloud is too sparse "
                 "for cutoff %g", cells_total, cutoff);
  } else {
    T width[3] = {T(0), T(0), T(0)};
    int64_t nc[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) {
      // ncell = clamp(floor(box / cutoff), 3), width = box / ncell, in T
      const T bt = (T)box[d];
      int64_t c = (int64_t)floor((double)(T)(bt / (T)cutoff));
      nc[d] = c < 3 ? 3 : c;
      width[d] = bt / (T)nc[d];
      cs.width[d] = (double)width[d];
      cs.ncell_per[d] = nc[d];
      cs.cmin[d] = 0;
      cs.ext[d] = nc[d];
    }
    k_cell_coords_periodic<T, D><<<grid, kBlock, 0, stream>>>(
        pos, n, width[0], width[1], width[2], nc[0], nc[1], nc[2], cs.cells.as<int64_t>());
    TSPH_CHECK_LAUNCH();
  }
  cs.stride[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) cs.stride[d] = cs.stride[d + 1] * cs.ext[d + 1];
  for (int d = D; d < 3; ++d) cs.stride[d] = 0;
  int axis_bits = 1;
  if (morton) {
    int64_t emax = 1;
    for (int d = 0; d < D; ++d) emax = cs.ext[d] > emax ? cs.ext[d] : emax;
    while (axis_bits < 31 && (1ll << axis_bits) < emax) ++axis_bits;
    TSPH_REQUIRE(axis_bits <= 63 / D, TSPH_ERR_ARG, "cell grid of %lld cells per axis exceeds the %d-bit "
                 "Morton key range", (long long)emax, 63 / D);
  }
  k_keys<D><<<grid, kBlock, 0, stream>>>(cs.cells.as<int64_t>(), n, cs.cmin[0], cs.cmin[1], cs.cmin[2],
                                         cs.stride[0], cs.stride[1], cs.stride[2], box ? 0 : 1,
                                         morton ? 1 : 0, cs.keys.as<int64_t>(), cs.idx_in.as<int32_t>());
  TSPH_CHECK_LAUNCH();
  // stable LSD radix sort over the key bits actually used
  double maxkey = 1.0;
  for (int d = 0; d < D; ++d) maxkey *= (double)cs.ext[d];
  int end_bit = 1;
  while (end_bit < 63 && (double)(1ull << end_bit) < maxkey) ++end_bit;
  if (morton) end_bit = D * axis_bits;
  size_t bytes = 0;
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(
      nullptr, bytes, (const unsigned long long*)cs.keys.ptr, (unsigned long long*)cs.skeys.ptr,
      cs.idx_in.as<int32_t>(), cs.order32.as<int32_t>(), (int)n, 0, end_bit, stream));
  TSPH_TRY(cs.sort_tmp.ensure(bytes));
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(
      cs.sort_tmp.ptr, bytes, (const unsigned long long*)cs.keys.ptr, (unsigned long long*)cs.skeys.ptr,
      cs.idx_in.as<int32_t>(), cs.order32.as<int32_t>(), (int)n, 0, end_bit, stream));
  int32_t* flags = cs.flags.as<int32_t>();
  int32_t* cid = flags + n;
  k_flag_starts<<<grid, kBlock, 0, stream>>>(cs.skeys.as<int64_t>(), n, flags, cs.order32.as<int32_t>(),
                                             cs.order.as<int64_t>());
  TSPH_CHECK_LAUNCH();
  bytes = 0;
  TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, flags, cid, (int)n, stream));
  TSPH_TRY(ctx->scan_tmp.ensure(bytes));
  TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(ctx->scan_tmp.ptr, bytes, flags, cid, (int)n, stream));
  k_write_cells<<<grid, kBlock, 0, stream>>>(cs.skeys.as<int64_t>(), flags, cid, n,
                                             cs.ukeys.as<int64_t>(), cs.ustart.as<int64_t>());
  TSPH_CHECK_LAUNCH();
  int32_t m = 0;
  TSPH_CHECK_HIP(hipMemcpyAsync(&m, cid + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  cs.ncells = m;
  return TSPH_OK;
}

template <typename T, int D>
int search_pairs(tsph_ctx* ctx, const T* pos, int64_t n, double cutoff, const double* box,
                 int64_t fluid_count, hipStream_t stream, int64_t* npairs) {
  ctx->npairs = 0;
  *npairs = 0;
  if (n < 2) { ctx->search.n = n; ctx->search.ncells = 0; return TSPH_OK; }
  TSPH_TRY((run_search<T, D>(ctx, pos, n, cutoff, box, stream)));
  CellSearch& cs = ctx->search;
  GridDesc g;
  for (int d = 0; d < 3; ++d) { g.ext[d] = cs.ext[d]; g.stride[d] = cs.stride[d]; }
  g.morton = cs.morton;
  const T c2 = (T)(cutoff * cutoff);
  T b[3] = {T(0), T(0), T(0)};
  if (box) for (int d = 0; d < D; ++d) b[d] = (T)box[d];
  TSPH_TRY(ctx->pair_count.ensure(sizeof(int64_t) * (n + 1)));
  TSPH_TRY(ctx->pair_off.ensure(sizeof(int64_t) * (n + 1)));
  const unsigned grid = grid_for(n);
  if (box)
    k_pairs<T, D, true, 0><<<grid, kBlock, 0, stream>>>(
        pos, n, cs.cells.as<int64_t>(), g, cs.ukeys.as<int64_t>(), cs.ustart.as<int64_t>(),
        cs.order.as<int64_t>(), cs.ncells, c2, b[0], b[1], b[2], fluid_count,
        ctx->pair_count.as<int64_t>(), nullptr, nullptr);
  else
    k_pairs<T, D, false, 0><<<grid, kBlock, 0, stream>>>(
        pos, n, cs.cells.as<int64_t>(), g, cs.ukeys.as<int64_t>(), cs.ustart.as<int64_t>(),
        cs.order.as<int64_t>(), cs.ncells, c2, b[0], b[1], b[2], fluid_count,
        ctx->pair_count.as<int64_t>(), nullptr, nullptr);
  TSPH_CHECK_LAUNCH();
  int64_t total = 0;
  TSPH_TRY(exclusive_scan(ctx, ctx->pair_count.as<unsigned long long>(), n,
                          ctx->pair_off.as<int64_t>(), stream, &total));
  ctx->npairs = total;
  *npairs = total;
  return TSPH_OK;
}

template <typename T, int D>
int fill_pairs(tsph_ctx* ctx, const T* pos, double cutoff, const double* box, int64_t fluid_count,
               int64_t* out, hipStream_t stream) {
  CellSearch& cs = ctx->search;
  const int64_t n = cs.n;
  GridDesc g;
  for (int d = 0; d < 3; ++d) { g.ext[d] = cs.ext[d]; g.stride[d] = cs.stride[d]; }
  g.morton = cs.morton;
  const T c2 = (T)(cutoff * cutoff);
  T b[3] = {T(0), T(0), T(0)};
  if (box) for (int d = 0; d < D; ++d) b[d] = (T)box[d];
  const unsigned grid = grid_for(n);
  if (box)
    k_pairs<T, D, true, 1><<<grid, kBlock, 0, stream>>>(
        pos, n, cs.cells.as<int64_t>(), g, cs.ukeys.as<int64_t>(), cs.ustart.as<int64_t>(),
        cs.order.as<int64_t>(), cs.ncells, c2, b[0], b[1], b[2], fluid_count, nullptr,
        ctx->pair_off.as<int64_t>(), out);
  else
    k_pairs<T, D, false, 1><<<grid, kBlock, 0, stream>>>(
        pos, n, cs.cells.as<int64_t>(), g, cs.ukeys.as<int64_t>(), cs.ustart.as<int64_t>(),
        cs.order.as<int64_t>(), cs.ncells, c2, b[0], b[1], b[2], fluid_count, nullptr,
        ctx->pair_off.as<int64_t>(), out);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// ---------------------------------------------------------------- brute force
// The O(N^2) search of brute_force_pairs (neighbor_search.py:29-41) and
// brute_force_pairs_min_image (neighbors3.py:145-170): every row tests every
// later row, with no cell machinery shared with the linked-cell search.  Rows
// are staged through LDS a block at a time; a block's threads start at the
// block's own first row, since earlier rows only pair with them as i < j.
// MODE 0 counts each row's partners j > i, MODE 1 writes them in ascending j.
template <typename T, int D, bool PER, int MODE>
__global__ void __launch_bounds__(kBlock) k_brute_pairs(const T* __restrict__ pos, int64_t n, T c2, T b0, T b1,
                                                         T b2, int64_t* __restrict__ count,
                                                         const int64_t* __restrict__ off,
                                                         int64_t* __restrict__ out) {
  __shared__ T tile[kBlock * D];
  const int64_t first = (int64_t)blockIdx.x * kBlock;
  const int64_t i = first + threadIdx.x;
  const T box[3] = {b0, b1, b2};
  T xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = i < n ? pos[i * D + d] : T(0);
  const int64_t base = (MODE == 1 && i < n) ? off[i] : 0;
  int64_t k = 0;
  for (int64_t j0 = first; j0 < n; j0 += kBlock) {
    __syncthreads();
    const int64_t jl = j0 + threadIdx.x;
#pragma unroll
    for (int d = 0; d < D; ++d) tile[threadIdx.x * D + d] = jl < n ? pos[jl * D + d] : T(0);
    __syncthreads();
    const int64_t m = n - j0 < kBlock ? n - j0 : kBlock;
    if (i >= n) continue;
    for (int64_t t = 0; t < m; ++t) {
      const int64_t j = j0 + t;
      if (j <= i) continue;
      if (dist2_pt<T, D, PER>(xi, tile + t * D, box) <= c2) {
        if (MODE == 1) {
          out[2 * (base + k)] = i;
          out[2 * (base + k) + 1] = j;
        }
        ++k;
      }
    }
  }
  if (MODE == 0 && i < n) count[i] = k;
}

template <typename T, int D>
int brute_pairs(tsph_ctx* ctx, const T* pos, int64_t n, double cutoff, const double* box, int mode,
                int64_t* out, hipStream_t stream, int64_t* npairs) {
  const T c2 = (T)(cutoff * cutoff);
  T b[3] = {T(0), T(0), T(0)};
  if (box) for (int d = 0; d < D; ++d) b[d] = (T)box[d];
  const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  if (mode == 0) {
    ctx->npairs = 0;
    *npairs = 0;
    if (n < 2) return TSPH_OK;
    TSPH_TRY(ctx->pair_count.ensure(sizeof(int64_t) * (n + 1)));
    TSPH_TRY(ctx->pair_off.ensure(sizeof(int64_t) * (n + 1)));
    if (box)
      k_brute_pairs<T, D, true, 0><<<grid, kBlock, 0, stream>>>(pos, n, c2, b[0], b[1], b[2],
                                                                 ctx->pair_count.as<int64_t>(), nullptr, nullptr);
    else
      k_brute_pairs<T, D, false, 0><<<grid, kBlock, 0, stream>>>(pos, n, c2, b[0], b[1], b[2],
                                                                  ctx->pair_count.as<int64_t>(), nullptr, nullptr);
    TSPH_CHECK_LAUNCH();
    int64_t total = 0;
    TSPH_TRY(exclusive_scan(ctx, ctx->pair_count.as<unsigned long long>(), n, ctx->pair_off.as<int64_t>(),
                            stream, &total));
    ctx->npairs = total;
    *npairs = total;
    return TSPH_OK;
  }
  if (box)
    k_brute_pairs<T, D, true, 1><<<grid, kBlock, 0, stream>>>(pos, n, c2, b[0], b[1], b[2], nullptr,
                                                               ctx->pair_off.as<int64_t>(), out);
  else
    k_brute_pairs<T, D, false, 1><<<grid, kBlock, 0, stream>>>(pos, n, c2, b[0], b[1], b[2], nullptr,
                                                                ctx->pair_off.as<int64_t>(), out);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// remembered between tsph_pairs_search and tsph_pairs_export
struct LastSearch {
  const void* pos = nullptr;
  int dtype = 0, dim = 0;
  double cutoff = 0.0;
  double box[3] = {0, 0, 0};
  int has_box = 0;
  int64_t fluid_count = -1;
  int brute = 0;
  int64_t n = 0;
};
static thread_local LastSearch g_last;

template <typename T> TSPH_D T absd(T x) { return x < T(0) ? -x : x; }

template <typename T, int D>
__global__ void k_max_disp(const T* __restrict__ pos, const T* __restrict__ anchor, int64_t n,
                           unsigned int* __restrict__ out_bits_f32, unsigned long long* __restrict__ out_bits_f64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  T v = T(0);
  if (i < n) {
    T acc = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T df = pos[i * D + d] - anchor[i * D + d];
      acc += df * df;
    }
    v = sqrt_s<T>(acc);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0) {
    // non-negative IEEE values order like their bit patterns; a wave whose
    // maximum does not beat the current value (a plain, possibly stale read)
    // skips its atomic, so the waves do not all serialise on the one word
    if (sizeof(T) == 4) {
      const unsigned int b = __float_as_uint((float)v);
      if (b > __atomic_load_n(out_bits_f32, __ATOMIC_RELAXED)) atomicMax(out_bits_f32, b);
    } else {
      const unsigned long long b = (unsigned long long)__double_as_longlong((double)v);
      if (b > __atomic_load_n(out_bits_f64, __ATOMIC_RELAXED)) atomicMax(out_bits_f64, b);
    }
  }
}

template int run_search<double, 2>(tsph_ctx*, const double*, int64_t, double, const double*, hipStream_t, bool);
template int run_search<double, 3>(tsph_ctx*, const double*, int64_t, double, const double*, hipStream_t, bool);

}  // namespace tsph

using namespace tsph;

extern "C" {

const char* tsph_last_error(void) { return g_err; }
int tsph_abi_version(void) { return TSPH_ABI_VERSION; }

int tsph_ctx_create(tsph_ctx** out) {
  TSPH_REQUIRE(out != nullptr, TSPH_ERR_ARG, "out must not be NULL");
  *out = new tsph_ctx();
  return TSPH_OK;
}

int tsph_ctx_destroy(tsph_ctx* ctx) {
  if (!ctx) return TSPH_OK;
  CellSearch& cs = ctx->search;
  DevBuf* bufs[] = {&cs.cells, &cs.keys, &cs.skeys, &cs.order, &cs.order32, &cs.ukeys, &cs.ustart,
                    &cs.flags, &cs.minmax, &cs.mm_blk, &cs.sort_tmp, &cs.idx_in, &ctx->pair_count, &ctx->pair_off,
                    &ctx->scan_tmp, &ctx->nb_off, &ctx->nb_idx, &ctx->nb_cursor, &ctx->wnb_off,
                    &ctx->wnb_idx, &ctx->wnb_cursor, &ctx->s_rho_raw, &ctx->s_rho, &ctx->s_p,
                    &ctx->s_aux, &ctx->s_aux2, &ctx->s_aux3, &ctx->s_gr, &ctx->s_flags, &ctx->s_pack, &ctx->s_red, &ctx->s_dp,
                    &ctx->s_four, &ctx->s_lap1, &ctx->s_cl, &ctx->g_part, &ctx->g_flag, &ctx->g_idx,
                    &ctx->g_tmp, &ctx->g_out};
  for (DevBuf* b : bufs) b->release();
  delete ctx;
  return TSPH_OK;
}

int tsph_pairs_search(tsph_ctx* ctx, const void* pos, int dtype, int64_t n, int dim, double cutoff,
                      const double* box, int64_t fluid_count, hipStream_t stream, int64_t* npairs) {
  TSPH_REQUIRE(ctx && npairs, TSPH_ERR_ARG, "ctx and npairs must not be NULL");
  TSPH_REQUIRE(dim >= 1 && dim <= 3, TSPH_ERR_ARG, "positions must have 1 to 3 dimensions, got %d", dim);
  TSPH_REQUIRE(dtype == TSPH_F64 || dtype == TSPH_F32, TSPH_ERR_TYPE, "positions must be float32 or float64");
  TSPH_REQUIRE(cutoff > 0.0 && cutoff < 1e300, TSPH_ERR_ARG, "cutoff must be finite and positive");
  TSPH_REQUIRE(n >= 0 && n < (1ll << 31) - 1, TSPH_ERR_ARG, "particle count %lld out of range", (long long)n);
  if (box) for (int d = 0; d < dim; ++d)
    TSPH_REQUIRE(box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive");
  g_last.pos = pos; g_last.dtype = dtype; g_last.dim = dim; g_last.cutoff = cutoff;
  g_last.has_box = box != nullptr; g_last.fluid_count = fluid_count; g_last.brute = 0; g_last.n = n;
  if (box) for (int d = 0; d < dim; ++d) g_last.box[d] = box[d];
#define TSPH_SEARCH(T, D) return search_pairs<T, D>(ctx, (const T*)pos, n, cutoff, box, fluid_count, stream, npairs)
  if (dtype == TSPH_F64) {
    if (dim == 1) TSPH_SEARCH(double, 1);
    if (dim == 2) TSPH_SEARCH(double, 2);
    TSPH_SEARCH(double, 3);
  } else {
    if (dim == 1) TSPH_SEARCH(float, 1);
    if (dim == 2) TSPH_SEARCH(float, 2);
    TSPH_SEARCH(float, 3);
  }
#undef TSPH_SEARCH
}

int tsph_pairs_brute(tsph_ctx* ctx, const void* pos, int dtype, int64_t n, int dim, double cutoff,
                     const double* box, hipStream_t stream, int64_t* npairs) {
  TSPH_REQUIRE(ctx && npairs, TSPH_ERR_ARG, "ctx and npairs must not be NULL");
  TSPH_REQUIRE(dim >= 1 && dim <= 3, TSPH_ERR_ARG, "positions must have 1 to 3 dimensions, got %d", dim);
  TSPH_REQUIRE(dtype == TSPH_F64 || dtype == TSPH_F32, TSPH_ERR_TYPE, "positions must be float32 or float64");
  TSPH_REQUIRE(cutoff > 0.0 && cutoff < 1e300, TSPH_ERR_ARG, "cutoff must be finite and positive");
  TSPH_REQUIRE(n >= 0 && n <= (1ll << 20), TSPH_ERR_ARG,
               "the O(N^2) search takes at most 2^20 particles, got %lld", (long long)n);
  if (box) for (int d = 0; d < dim; ++d)
    TSPH_REQUIRE(box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive");
  g_last.pos = pos; g_last.dtype = dtype; g_last.dim = dim; g_last.cutoff = cutoff;
  g_last.has_box = box != nullptr; g_last.fluid_count = -1; g_last.brute = 1; g_last.n = n;
  if (box) for (int d = 0; d < dim; ++d) g_last.box[d] = box[d];
#define TSPH_BRUTE(T, D) return brute_pairs<T, D>(ctx, (const T*)pos, n, cutoff, box, 0, nullptr, stream, npairs)
  if (dtype == TSPH_F64) {
    if (dim == 1) TSPH_BRUTE(double, 1);
    if (dim == 2) TSPH_BRUTE(double, 2);
    TSPH_BRUTE(double, 3);
  } else {
    if (dim == 1) TSPH_BRUTE(float, 1);
    if (dim == 2) TSPH_BRUTE(float, 2);
    TSPH_BRUTE(float, 3);
  }
#undef TSPH_BRUTE
}

int tsph_pairs_export(tsph_ctx* ctx, int64_t* pairs, hipStream_t stream) {
  TSPH_REQUIRE(ctx, TSPH_ERR_ARG, "ctx must not be NULL");
  if (ctx->npairs == 0) return TSPH_OK;
  TSPH_REQUIRE(pairs, TSPH_ERR_ARG, "pairs must not be NULL");
  const double* box = g_last.has_box ? g_last.box : nullptr;
  if (g_last.brute) {
    int64_t unused = 0;
#define TSPH_BRUTE(T, D) return brute_pairs<T, D>(ctx, (const T*)g_last.pos, g_last.n, g_last.cutoff, box, 1, pairs, stream, &unused)
    if (g_last.dtype == TSPH_F64) {
      if (g_last.dim == 1) TSPH_BRUTE(double, 1);
      if (g_last.dim == 2) TSPH_BRUTE(double, 2);
      TSPH_BRUTE(double, 3);
    } else {
      if (g_last.dim == 1) TSPH_BRUTE(float, 1);
      if (g_last.dim == 2) TSPH_BRUTE(float, 2);
      TSPH_BRUTE(float, 3);
    }
#undef TSPH_BRUTE
  }
#define TSPH_FILL(T, D) return fill_pairs<T, D>(ctx, (const T*)g_last.pos, g_last.cutoff, box, g_last.fluid_count, pairs, stream)
  if (g_last.dtype == TSPH_F64) {
    if (g_last.dim == 1) TSPH_FILL(double, 1);
    if (g_last.dim == 2) TSPH_FILL(double, 2);
    TSPH_FILL(double, 3);
  } else {
    if (g_last.dim == 1) TSPH_FILL(float, 1);
    if (g_last.dim == 2) TSPH_FILL(float, 2);
    TSPH_FILL(float, 3);
  }
#undef TSPH_FILL
}

int tsph_cells_export(tsph_ctx* ctx, int64_t* cells, int64_t* keys, int64_t* order, int64_t* ncells,
                      hipStream_t stream) {
  TSPH_REQUIRE(ctx, TSPH_ERR_ARG, "ctx must not be NULL");
  CellSearch& cs = ctx->search;
  if (ncells) *ncells = cs.ncells;
  if (cs.n == 0 || cs.ncells == 0) return TSPH_OK;
  if (cells) TSPH_CHECK_HIP(hipMemcpyAsync(cells, cs.cells.ptr, sizeof(int64_t) * cs.n * cs.dim,
                                           hipMemcpyDeviceToDevice, stream));
  if (keys) TSPH_CHECK_HIP(hipMemcpyAsync(keys, cs.keys.ptr, sizeof(int64_t) * cs.n,
                                          hipMemcpyDeviceToDevice, stream));
  if (order) TSPH_CHECK_HIP(hipMemcpyAsync(order, cs.order.ptr, sizeof(int64_t) * cs.n,
                                           hipMemcpyDeviceToDevice, stream));
  return TSPH_OK;
}

int tsph_max_displacement(const void* pos, const void* anchor, int dtype, int64_t n, int dim, void* out,
                          hipStream_t stream) {
  TSPH_REQUIRE(out, TSPH_ERR_ARG, "out must not be NULL");
  TSPH_REQUIRE(dim >= 1 && dim <= 3, TSPH_ERR_ARG, "dim must be 1..3");
  const size_t sz = dtype == TSPH_F64 ? 8 : 4;
  TSPH_CHECK_HIP(hipMemsetAsync(out, 0, sz, stream));
  if (n == 0) return TSPH_OK;
  const unsigned grid = grid_for(n);
#define TSPH_DISP(T, D) k_max_disp<T, D><<<grid, kBlock, 0, stream>>>((const T*)pos, (const T*)anchor, n, (unsigned int*)out, (unsigned long long*)out)
  if (dtype == TSPH_F64) {
    if (dim == 1) TSPH_DISP(double, 1); else if (dim == 2) TSPH_DISP(double, 2); else TSPH_DISP(double, 3);
  } else {
    if (dim == 1) TSPH_DISP(float, 1); else if (dim == 2) TSPH_DISP(float, 2); else TSPH_DISP(float, 3);
  }
#undef TSPH_DISP
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

int tsph_nlist_from_pairs(tsph_ctx* ctx, const int64_t* pairs, int64_t npairs, int64_t n_fluid,
                          int64_t n_total, int all_targets, hipStream_t stream) {
  TSPH_REQUIRE(ctx, TSPH_ERR_ARG, "ctx must not be NULL");
  TSPH_REQUIRE(n_fluid >= 0 && n_total >= n_fluid && n_total < (1ll << 31) - 1, TSPH_ERR_ARG,
               "bad particle counts");
  TSPH_REQUIRE(npairs >= 0, TSPH_ERR_ARG, "npairs must be non-negative");
  const int64_t nt = all_targets ? n_total : n_fluid;
  const int64_t nw = all_targets ? 0 : n_total - n_fluid;
  ctx->nb_targets = nt; ctx->wnb_targets = nw;
  TSPH_TRY(ctx->nb_off.ensure(sizeof(int64_t) * (nt + 1)));
  TSPH_TRY(ctx->nb_cursor.ensure(sizeof(int64_t) * (nt + 1)));
  TSPH_TRY(ctx->wnb_off.ensure(sizeof(int64_t) * (nw + 1)));
  TSPH_TRY(ctx->wnb_cursor.ensure(sizeof(int64_t) * (nw + 1) + sizeof(int)));
  unsigned long long* cnt = ctx->nb_cursor.as<unsigned long long>();
  unsigned long long* wcnt = ctx->wnb_cursor.as<unsigned long long>();
  int* bad = (int*)(wcnt + nw + 1);
  TSPH_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int64_t) * (nt + 1), stream));
  TSPH_CHECK_HIP(hipMemsetAsync(wcnt, 0, sizeof(int64_t) * (nw + 1) + sizeof(int), stream));
  if (npairs > 0)
    k_count_pairs<<<grid_for(npairs), kBlock, 0, stream>>>(pairs, npairs, n_fluid, n_total,
                                                            all_targets, cnt, wcnt, bad);
  TSPH_CHECK_LAUNCH();
  int hbad = 0;
  TSPH_CHECK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, stream));
  TSPH_TRY(exclusive_scan(ctx, cnt, nt, ctx->nb_off.as<int64_t>(), stream, &ctx->nb_entries));
  TSPH_REQUIRE(hbad == 0, TSPH_ERR_ARG,
               "pairs must hold indices 0 <= i, j < %lld with i != j", (long long)n_total);
  TSPH_TRY(exclusive_scan(ctx, wcnt, nw, ctx->wnb_off.as<int64_t>(), stream, &ctx->wnb_entries));
  TSPH_TRY(ctx->nb_idx.ensure(sizeof(int32_t) * (ctx->nb_entries + 1)));
  TSPH_TRY(ctx->wnb_idx.ensure(sizeof(int32_t) * (ctx->wnb_entries + 1)));
  TSPH_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int64_t) * (nt + 1), stream));
  TSPH_CHECK_HIP(hipMemsetAsync(wcnt, 0, sizeof(int64_t) * (nw + 1), stream));
  if (npairs > 0)
    k_fill_pairs<<<grid_for(npairs), kBlock, 0, stream>>>(
        pairs, npairs, n_fluid, all_targets, ctx->nb_off.as<int64_t>(), cnt, ctx->nb_idx.as<int32_t>(),
        ctx->wnb_off.as<int64_t>(), wcnt, ctx->wnb_idx.as<int32_t>());
  TSPH_CHECK_LAUNCH();
  if (nt > 0) k_sort_lists<<<grid_for(nt), kBlock, 0, stream>>>(ctx->nb_off.as<int64_t>(), nt,
                                                                   ctx->nb_idx.as<int32_t>());
  if (nw > 0) k_sort_lists<<<grid_for(nw), kBlock, 0, stream>>>(ctx->wnb_off.as<int64_t>(), nw,
                                                                   ctx->wnb_idx.as<int32_t>());
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

}  // extern "C"
