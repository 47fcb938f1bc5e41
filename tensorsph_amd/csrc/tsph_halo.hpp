// Neighbourhood halos of the 2-D resident stepper's slot-list sweeps.
//
// The rows of a workgroup (kHaloRows consecutive fluid rows, a tile) are
// Morton-adjacent, so their neighbour lists name a small set of rows: those
// of the 3^D cells around the tile's cells.  At each list rebuild a tile's
// halo is built once: the union of those cells in Morton order (which is the
// order of the rows), as the explicit list of its fluid rows followed by its
// wall rows, and every list entry of the tile's rows re-expressed as a 16-bit
// position in that halo.  A sweep then stages the halo's rows in LDS with
// mostly contiguous loads and evaluates its lists from LDS.  The lists keep
// their order and their entries name the same rows, so a sweep gives the same
// bits from the halo as from the global arrays.  A tile whose neighbourhood
// is too wide (sparse spray, more than kHaloCand candidate cells or more rows
// than the sweeps' LDS holds) keeps the global path (count -1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsph_cells.hpp"
#include "tsph_common.hpp"

namespace tsph {

constexpr int kHaloRows = 256;     // rows per tile: the sweeps' workgroup
constexpr int kHaloCand = 1024;    // neighbourhood cells a tile may name
constexpr int kHaloCap = 1024;     // halo rows a tile may stage

// Per tile: halo rows, fluid halo rows first (x = all, y = fluid; x = -1: the
// tile keeps the global path).
struct HaloView {
  const uint32_t* rows;            // (tiles, kHaloCap)
  const int2* cnt;
  const uint16_t* idx;             // lists of halo positions, NbSlots layout
  int64_t ntiles;
};

// f(l4, keep4) once per chunk of four entries of row i's halo-position list
// (the NbSlots layout with 16-bit entries): padded slots hold position 0
// with keep false.  CH chunks per batch: the batch's loads come first.
struct NbHalo {
  const uint16_t* idx;
  const int32_t* cnt;
  int64_t K;
  template <int D, int CH = 1, typename F>
  TSPH_D void visit_chunks(int64_t i, F&& f) const {
    const ushort4* p = reinterpret_cast<const ushort4*>(idx) + (i >> 6) * 16 * K + (i & 63);
    const int n = cnt[i];
    for (int k = 0; k < n; k += 4 * CH) {
      ushort4 q[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) q[c] = k + 4 * c < n ? p[((k >> 2) + c) * 64] : make_ushort4(0, 0, 0, 0);
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int32_t raw[4] = {q[c].x, q[c].y, q[c].z, q[c].w};
        int32_t l4[4];
        bool keep4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool in = k + 4 * c + u < n;
          l4[u] = in ? raw[u] : 0;
          keep4[u] = in;
        }
        f(l4, keep4);
      }
    }
  }
};

// exclusive prefix of one value per thread of a 256-thread block; wtot is
// 4 ints of LDS; the block total comes back in *total
TSPH_D int halo_block_scan(int v, int* wtot, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  int base = 0;
  for (int k = 0; k < w; ++k) base += wtot[k];
  *total = wtot[0] + wtot[1] + wtot[2] + wtot[3];
  __syncthreads();
  return base + inc - v;
}

// prefix over arr[0, m) (m <= 4 * 256) in place, four consecutive entries per
// thread; returns the total
TSPH_D int halo_array_scan(int* arr, int m, int* wtot) {
  const int t = threadIdx.x;
  int loc[4], s = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = 4 * t + q;
    loc[q] = e < m ? arr[e] : 0;
    s += loc[q];
  }
  int total;
  int base = halo_block_scan(s, wtot, &total);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = 4 * t + q;
    if (e < m) arr[e] = base;
    base += loc[q];
  }
  __syncthreads();
  return total;
}

TSPH_D uint64_t compact_bits2(uint64_t v) {
  v &= 0x5555555555555555ull;
  v = (v | (v >> 1)) & 0x3333333333333333ull;
  v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
  v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
  v = (v | (v >> 16)) & 0x00000000ffffffffull;
  return v;
}

// the largest q in [0, m) with a[q] <= x (a ascending, a[0] <= x)
TSPH_D int halo_floor_search(const int* a, int m, int x) {
  int lo = 0, hi = m;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// One workgroup per tile (2-D).  cg: the cell table of the build (cells of
// every row, per-cell fluid / wall row ranges); nb: the tile rows' global
// lists; cap (<= kHaloCap): the most rows a halo may have.  Writes the halo
// rows, the counts and the halo-position lists; hmax[0] gets the largest
// halo and hmax[1] counts the tiles on the halo path.
__global__ void __launch_bounds__(kHaloRows) k_halo_build2(NbCells cg, NbSlots nb, int64_t n, int cap,
                                                           uint32_t* __restrict__ halo_rows,
                                                           int2* __restrict__ halo_cnt,
                                                           uint16_t* __restrict__ lidx,
                                                           unsigned int* __restrict__ hmax) {
  constexpr int D = 2, NOFF = 9;
  __shared__ unsigned long long key[kHaloCand];
  __shared__ int fstart[kHaloCand], foff[kHaloCand], wstart[kHaloCand], woff[kHaloCand];
  __shared__ int flen[kHaloCand], wlen[kHaloCand];
  __shared__ int wtot[4];
  __shared__ int s_bad;
  const int t = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int64_t r = tile * kHaloRows + t;
  const bool live = r < n;
  int64_t c[2] = {0, 0};
  bool head = false;
  if (live) {
    c[0] = cg.cells[r * D];
    c[1] = cg.cells[r * D + 1];
    head = t == 0 || cg.cells[(r - 1) * D] != c[0] || cg.cells[(r - 1) * D + 1] != c[1];
  }
  if (t == 0) s_bad = 0;
  int nheads;
  const int hpos = halo_block_scan(head ? 1 : 0, wtot, &nheads);
  const int ncand = nheads * NOFF;
  if (ncand > kHaloCand) {                  // too many distinct cells: global path
    if (t == 0) halo_cnt[tile] = make_int2(-1, -1);
    return;
  }
  int np = 1;
  while (np < ncand) np <<= 1;
  // the 3^D cells around every distinct cell of the tile, as Morton keys
  if (head) {
    for (int o = 0; o < NOFF; ++o) {
      const int64_t nc[2] = {c[0] + (o / 3 - 1), c[1] + (o % 3 - 1)};
      const bool in = nc[0] >= 0 && nc[0] < cg.ext[0] && nc[1] >= 0 && nc[1] < cg.ext[1];
      key[hpos * NOFF + o] = in ? (unsigned long long)morton_key<2>(nc) : ~0ull;
    }
  }
  for (int s = ncand + t; s < np; s += kHaloRows) key[s] = ~0ull;
  __syncthreads();
  // bitonic sort of key[0, np)
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int p = t; p < (np >> 1); p += kHaloRows) {
        const int lo = 2 * p - (p & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned long long a = key[lo], b = key[hi];
        if ((a > b) == up) { key[lo] = b; key[hi] = a; }
      }
      __syncthreads();
    }
  }
  // distinct cells, in Morton order: their fluid and wall row ranges
  int nu = 0;
  {
    int flag[4], s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q;
      flag[q] = e < np && key[e] != ~0ull && (e == 0 || key[e] != key[e - 1]) ? 1 : 0;
      s += flag[q];
    }
    int base = halo_block_scan(s, wtot, &nu);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q;
      if (flag[q]) {
        const unsigned long long k = key[e];
        const int64_t cx = (int64_t)compact_bits2(k >> 1), cy = (int64_t)compact_bits2(k);
        const int4 R = cg.R[cx * cg.stride[0] + cy * cg.stride[1]];
        fstart[base] = R.x; flen[base] = R.y - R.x;
        wstart[base] = R.z; wlen[base] = R.w - R.z;
        ++base;
      }
    }
  }
  __syncthreads();
  // keep the non-empty ranges (starts ascending with the Morton order)
  int nf = 0, nw = 0;
  {
    int ff[4], fw[4], sf = 0, sw = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q;
      ff[q] = e < nu && flen[e] > 0 ? 1 : 0;
      fw[q] = e < nu && wlen[e] > 0 ? 1 : 0;
      sf += ff[q]; sw += fw[q];
    }
    int bf = halo_block_scan(sf, wtot, &nf);
    int bw = halo_block_scan(sw, wtot, &nw);
    int fs[4], fl[4], ws[4], wl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q;
      fs[q] = e < nu ? fstart[e] : 0; fl[q] = e < nu ? flen[e] : 0;
      ws[q] = e < nu ? wstart[e] : 0; wl[q] = e < nu ? wlen[e] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (ff[q]) { fstart[bf] = fs[q]; foff[bf] = fl[q]; ++bf; }
      if (fw[q]) { wstart[bw] = ws[q]; woff[bw] = wl[q]; ++bw; }
    }
    __syncthreads();
  }
  const int F = halo_array_scan(foff, nf, wtot);     // lengths -> halo offsets
  const int W = halo_array_scan(woff, nw, wtot);
  const int H = F + W;
  if (H > cap) {
    if (t == 0) halo_cnt[tile] = make_int2(-1, -1);
    return;
  }
  // the halo's rows: fluid ranges, then wall ranges
  uint32_t* hr = halo_rows + tile * kHaloCap;
  for (int q = t; q < nf; q += kHaloRows) {
    const int len = (q + 1 < nf ? foff[q + 1] : F) - foff[q];
    for (int k = 0; k < len; ++k) hr[foff[q] + k] = (uint32_t)(fstart[q] + k);
  }
  for (int q = t; q < nw; q += kHaloRows) {
    const int len = (q + 1 < nw ? woff[q + 1] : W) - woff[q];
    for (int k = 0; k < len; ++k) hr[F + woff[q] + k] = (uint32_t)(wstart[q] + k);
  }
  // each row's entries as halo positions, chunk by chunk; an entry outside
  // the halo (a list built with a cutoff above the cell edge) sends the tile
  // to the global path
  bool bad = false;
  if (live) {
    const int4* src = nb.chunks(r);
    ushort4* dst = reinterpret_cast<ushort4*>(lidx) + (r >> 6) * 16 * nb.K + (r & 63);
    const int cntr = nb.cnt[r];
    for (int k = 0; k < cntr; k += 4) {
      const int4 qj = src[(k >> 2) * 64];
      const int32_t j4[4] = {qj.x, qj.y, qj.z, qj.w};
      uint16_t l4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int l = 0;
        if (k + u < cntr) {
          const int j = j4[u];
          const bool fluid = j < n;
          const int* st = fluid ? fstart : wstart;
          const int* of = fluid ? foff : woff;
          const int m = fluid ? nf : nw;
          const int end = fluid ? F : W;
          if (m == 0 || j < st[0]) {
            bad = true;
          } else {
            const int q = halo_floor_search(st, m, j);
            const int len = (q + 1 < m ? of[q + 1] : end) - of[q];
            if (j - st[q] >= len) bad = true;
            l = (fluid ? 0 : F) + of[q] + (j - st[q]);
          }
        }
        l4[u] = (uint16_t)(bad ? 0 : l);
      }
      dst[(k >> 2) * 64] = make_ushort4(l4[0], l4[1], l4[2], l4[3]);
    }
  }
  if (bad) s_bad = 1;
  __syncthreads();
  if (t == 0) {
    if (s_bad) {
      halo_cnt[tile] = make_int2(-1, -1);
    } else {
      halo_cnt[tile] = make_int2(H, F);
      atomicMax(hmax, (unsigned int)H);
      atomicAdd(hmax + 1, 1u);
    }
  }
}

}  // namespace tsph
