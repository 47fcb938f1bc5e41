// Neighbourhood halos of the resident stepper's slot-list sweeps (2-D and 3-D).
//
// The rows of a workgroup (kHaloRows consecutive fluid rows, a tile) are
// adjacent in the binning's blocked Morton order, so their neighbour lists
// name a small set of rows: those of the 3^D cells around the tile's cells.
// At each list rebuild a tile's halo is built once: the union of those
// cells, as the explicit list of its fluid rows followed by its wall rows,
// and every list entry of the tile's rows re-expressed as a 16-bit
// position in that halo.  A sweep then stages the halo's rows in LDS with
// mostly contiguous loads and evaluates its lists from LDS.  The lists keep
// their order and their entries name the same rows, so a sweep gives the same
// bits from the halo as from the global arrays.  A tile whose neighbourhood
// is too wide (sparse spray or a tile split across two regions: more than
// kHaloBox cells in the box around its cells, more than kHaloUniq
// neighbourhood cells, or more rows than the sweeps' LDS holds) keeps the
// global path (count -1).
#pragma once
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "tsph_bin.hpp"
#include "tsph_cells.hpp"
#include "tsph_common.hpp"

namespace tsph {

constexpr int kHaloRows = 256;     // rows per tile: the sweeps' workgroup
// halo rows a tile may stage: about 3x a 2-D tile's rows, 7-10x a 3-D tile's
// (the 27 cells around each of a tile's ten or so cells, ~25 rows a cell)
template <int D> constexpr int halo_cap() { return D == 2 ? 1024 : 2816; }

// Per tile: halo rows, fluid halo rows first (x = all, y = fluid; x = -1: the
// tile keeps the global path).
struct HaloView {
  const uint32_t* rows;            // (tiles, halo_cap<D>())
  const int2* cnt;
  const uint16_t* idx;             // lists of halo positions, NbSlots layout
  int64_t ntiles;
  int max_rows;                    // rows the launch stages (larger halos: global path)
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

// the largest q in [0, m) with a[q] <= x (a ascending, a[0] <= x)
TSPH_D int halo_floor_search(const int* a, int m, int x) {
  int lo = 0, hi = m;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

constexpr int kHaloUniq = 512;     // distinct neighbourhood cells a tile may hold
constexpr int kHaloBox = 1024;     // cells of the box around a tile's cells

// relative positions of the builder's halo rows in LDS
template <int D> struct HaloPosT;
template <> struct HaloPosT<2> { using T = float2; };
template <> struct HaloPosT<3> { using T = float4; };

// One workgroup per tile of kHaloRows fluid rows: the tile rows' slot lists
// at the build radius (rc2) in both forms, global rows (idx, the NbSlots
// layout) and halo positions (lidx), and the tile's halo.
//
// The tile's distinct cells are the runs of equal cells among its rows
// (sorted by binning key).  Each marks the 3^D cells around it in a table
// over the box of the tile's cells widened by one cell; the marked cells,
// numbered in box order, are the neighbourhood, and every (cell, offset)
// candidate reads its cell's number u from the table.  The halo holds the
// fluid rows of those cells in that order, then their wall rows, and their
// positions are staged in LDS.  A row then walks the cells of its own
// cell's 3^D candidates, in the cell walk's order (offsets with the last axis
// fastest, fluid rows before wall rows), tests each row by the squared
// distance (an FP32 test on tile-relative positions, the FP64 test within
// its rounding band: the test the other list builds decide by) and appends
// the accepted rows to both lists.  A tile with too many cells or rows for
// its LDS walks the global cell table instead and gets no halo (count -1).
//
// cnt, idx: per-row count and global lists (K slots); lidx: halo lists, the
// same layout with 16-bit entries; blk[2 * tile]: the tile's longest list and
// (halo rows + 1, or 0 without a halo).
//
// MAP: the global lists are already built (the cell walk, k_slot_lists) and
// only the halo and the halo lists are made: every entry j of a tile row's
// list lies in the 3^D cells around the row's cell, so its halo position is
// its cell's halo offset plus its place in the cell's fluid or wall run
// (blk[2 * tile] is 0: the walk reports the longest list).  The 3-D build
// takes this form: the walk's FP32 prefilter over the global positions
// fills the 3-D lists in about half the time of the staged-candidate walk.
template <int D, bool FILL, bool MAP = false>
__global__ void __launch_bounds__(kHaloRows) k_halo_lists(NbCells cg, const double* __restrict__ pos,
                                                          int64_t n, double rc2, int cap,
                                                          int32_t* __restrict__ cnt, int32_t* __restrict__ idx,
                                                          uint16_t* __restrict__ lidx, int64_t K,
                                                          uint32_t* __restrict__ halo_rows,
                                                          int2* __restrict__ halo_cnt,
                                                          unsigned int* __restrict__ blk) {
  constexpr int NOFF = D == 2 ? 9 : 27;
  constexpr int CAP = halo_cap<D>();
  using HP = typename HaloPosT<D>::T;
  __shared__ int16_t tbl[kHaloBox];                    // the tile's box of cells: -1, 0 (marked), or the cell's number
  __shared__ int16_t cand_u[kHaloRows * NOFF];
  __shared__ int ufs[kHaloUniq], ufl[kHaloUniq], ufo[kHaloUniq], uws[kHaloUniq], uwl[kHaloUniq], uwo[kHaloUniq];
  __shared__ HP hx[CAP];                               // halo positions relative to the tile's first row
  __shared__ int wtot[4];
  __shared__ int s_box[2 * D][kHaloRows / 64];
  __shared__ int s_kmax[4];
  __shared__ float s_ext[4];
  const int t = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int64_t r = tile * kHaloRows + t;
  const bool live = r < n;
  int c[D];
  bool head = false;
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = live ? (int)cg.cells[r * D + d] : 0;
  if (live) {
    head = t == 0;
#pragma unroll
    for (int d = 0; d < D; ++d) head = head || cg.cells[(r - 1) * D + d] != c[d];
  }
  int nheads;
  const int hpos = halo_block_scan(head ? 1 : 0, wtot, &nheads);
  const int hidx = hpos + (head ? 1 : 0) - 1;          // this row's cell among the tile's
  // the box of the tile's cells widened by one cell, cut to the grid
  int bb[2 * D];
#pragma unroll
  for (int d = 0; d < D; ++d) { bb[d] = live ? c[d] : INT_MAX; bb[D + d] = live ? c[d] : INT_MIN; }
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < 2 * D; ++q) {
      const int y = __shfl_xor(bb[q], o, 64);
      bb[q] = q < D ? (y < bb[q] ? y : bb[q]) : (y > bb[q] ? y : bb[q]);
    }
  }
  if ((t & 63) == 0)
    for (int q = 0; q < 2 * D; ++q) s_box[q][t >> 6] = bb[q];
  __syncthreads();
  for (int w = 0; w < kHaloRows / 64; ++w)
#pragma unroll
    for (int q = 0; q < 2 * D; ++q) {
      const int y = s_box[q][w];
      bb[q] = q < D ? (y < bb[q] ? y : bb[q]) : (y > bb[q] ? y : bb[q]);
    }
  int b0[D], be[D];                                    // box origin and edge lengths (cells)
  int64_t nbox = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    b0[d] = bb[d] - 1 > 0 ? bb[d] - 1 : 0;
    const int top = bb[D + d] + 1 < (int)cg.ext[d] - 1 ? bb[D + d] + 1 : (int)cg.ext[d] - 1;
    be[d] = top - b0[d] + 1;
    nbox *= be[d];
  }
  // box number of cell c + offset o (o: last axis fastest, as the cell walk),
  // or -1 outside the box
  auto box_of = [&](const int* cc, int o) -> int {
    int e = 0, rem = o;
    bool in = true;
    int off[D];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) { off[d] = rem % 3 - 1; rem /= 3; }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int v = cc[d] + off[d] - b0[d];
      in = in && v >= 0 && v < be[d];
      e = e * be[d] + v;
    }
    return in ? e : -1;
  };
  bool ok = nbox <= kHaloBox;                          // block-uniform
  int F = 0, W = 0;
  if (ok) {
    for (int e = t; e < kHaloBox; e += kHaloRows) tbl[e] = -1;
    __syncthreads();
    if (head) {
      for (int o = 0; o < NOFF; ++o) {
        const int e = box_of(c, o);
        if (e >= 0) tbl[e] = 0;
      }
    }
    __syncthreads();
    // number the marked cells (box order) and fetch their row ranges
    int flag[4], fsum = 0, nu = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = 4 * t + q;
      flag[q] = e < (int)nbox && tbl[e] == 0 ? 1 : 0;
      fsum += flag[q];
    }
    const int ubase = halo_block_scan(fsum, wtot, &nu);
    ok = nu <= kHaloUniq;
    if (ok) {
      int u = ubase;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!flag[q]) continue;
        const int e = 4 * t + q;
        tbl[e] = (int16_t)u;
        int64_t L = 0;
        int rem = e;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
          L += (int64_t)(b0[d] + rem % be[d]) * cg.stride[d];
          rem /= be[d];
        }
        const int4 R = cg.R[L];
        ufs[u] = R.x; ufl[u] = R.y - R.x; ufo[u] = R.y - R.x;
        uws[u] = R.z; uwl[u] = R.w - R.z; uwo[u] = R.w - R.z;
        ++u;
      }
    }
    __syncthreads();
    if (ok && head) {
      for (int o = 0; o < NOFF; ++o) {
        const int e = box_of(c, o);
        cand_u[hpos * NOFF + o] = e >= 0 ? tbl[e] : (int16_t)-1;
      }
    }
    if (ok) {
      F = halo_array_scan(ufo, nu, wtot);               // lengths -> halo offsets
      W = halo_array_scan(uwo, nu, wtot);
      ok = F + W <= cap;
    }
    if (ok) {
      // the halo's rows and (unless MAP) their positions: fluid rows by cell, then wall
      // rows; slot s of the halo finds its cell by the offsets, and every
      // thread issues its (up to CAP / kHaloRows) loads together
      uint32_t* hr = halo_rows + tile * CAP;
      const int H = F + W;
      constexpr int QS = CAP / kHaloRows;
      int32_t jq[QS];
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s2 = t + q * kHaloRows;
        int32_t j = 0;
        if (s2 < H) {
          const bool wl = s2 >= F;
          const int x = wl ? s2 - F : s2;
          const int* off = wl ? uwo : ufo;
          const int* len = wl ? uwl : ufl;
          const int* st = wl ? uws : ufs;
          // the last cell whose offset is <= x (the one holding slot x)
          int lo = 0, hi = nu;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= x) lo = mid; else hi = mid;
          }
          while (len[lo] == 0) --lo;
          j = st[lo] + (x - off[lo]);
          hr[s2] = (uint32_t)j;
        }
        jq[q] = j;
      }
      if constexpr (!MAP) {
      double org[D];
#pragma unroll
      for (int d = 0; d < D; ++d) org[d] = pos[tile * kHaloRows * D + d];
      double xq[QS][D];
#pragma unroll
      for (int q = 0; q < QS; ++q)
#pragma unroll
        for (int d = 0; d < D; ++d) xq[q][d] = pos[(int64_t)jq[q] * D + d];
      float ext = 0.f;
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s2 = t + q * kHaloRows;
        if (s2 < H) {
          float v[D];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            v[d] = (float)(xq[q][d] - org[d]);
            ext = fmaxf(ext, fabsf(v[d]));
          }
          if constexpr (D == 2) hx[s2] = make_float2(v[0], v[1]);
          else hx[s2] = make_float4(v[0], v[1], v[2], 0.f);
        }
      }
      for (int o = 32; o > 0; o >>= 1) ext = fmaxf(ext, __shfl_xor(ext, o, 64));
      if ((t & 63) == 0) s_ext[t >> 6] = ext;
      __syncthreads();
      }
    }
  }
  if constexpr (MAP) {
    // the halo positions of the global lists' entries
    if (live && ok) {
      const int nr = cnt[r];
      const int4* gp = reinterpret_cast<const int4*>(idx) + (r >> 6) * 16 * K + (r & 63);
      ushort4* hp = reinterpret_cast<ushort4*>(lidx) + (r >> 6) * 16 * K + (r & 63);
      for (int k0 = 0; k0 < nr; k0 += 4) {
        const int4 q = gp[(k0 >> 2) * 64];
        const int32_t j4[4] = {q.x, q.y, q.z, q.w};
        uint16_t l4[4];
#pragma unroll
        for (int uu = 0; uu < 4; ++uu) {
          int l = 0;
          if (k0 + uu < nr) {
            const int32_t j = j4[uu];
            int cj[D];
#pragma unroll
            for (int d = 0; d < D; ++d) cj[d] = (int)cg.cells[(int64_t)j * D + d];
            const int e = box_of(cj, NOFF / 2);            // offset NOFF / 2: the cell itself
            const int u = e >= 0 ? tbl[e] : -1;            // always a marked cell (the walk's)
            if (u >= 0) l = j < n ? ufo[u] + (j - ufs[u]) : F + uwo[u] + (j - uws[u]);
          }
          l4[uu] = (uint16_t)l;
        }
        hp[(k0 >> 2) * 64] = make_ushort4(l4[0], l4[1], l4[2], l4[3]);
      }
    }
    if (t == 0) {
      blk[2 * tile] = 0u;
      blk[2 * tile + 1] = ok ? (unsigned int)(F + W + 1) : 0u;
      halo_cnt[tile] = ok ? make_int2(F + W, F) : make_int2(-1, -1);
    }
    return;
  }
  // the lists
  int k = 0;
  if (live) {
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = pos[r * D + d];
    int4* gp = FILL ? reinterpret_cast<int4*>(idx) + (r >> 6) * 16 * K + (r & 63) : nullptr;
    ushort4* hp = FILL ? reinterpret_cast<ushort4*>(lidx) + (r >> 6) * 16 * K + (r & 63) : nullptr;
    const int chunks = (int)(K >> 2);
    int4 gb = make_int4((int)r, (int)r, (int)r, (int)r);
    ushort4 hb = make_ushort4(0, 0, 0, 0);
    auto accept = [&](int32_t j, int l) {
      if (FILL) {
        const int sl = k & 3;
        gb.x = sl == 0 ? j : gb.x; gb.y = sl == 1 ? j : gb.y; gb.z = sl == 2 ? j : gb.z; gb.w = sl == 3 ? j : gb.w;
        const uint16_t l16 = (uint16_t)l;
        hb.x = sl == 0 ? l16 : hb.x; hb.y = sl == 1 ? l16 : hb.y; hb.z = sl == 2 ? l16 : hb.z; hb.w = sl == 3 ? l16 : hb.w;
        if (sl == 3 && (k >> 2) < chunks) {
          gp[(k >> 2) * 64] = gb;
          if (ok) hp[(k >> 2) * 64] = hb;
        }
      }
      ++k;
    };
    if (ok) {
      // FP32 test on the tile-relative positions: below lo2 accepts, above
      // hi2 rejects, the band between (which holds the FP32 rounding error of
      // the test) is decided by the FP64 test on the global positions.  The
      // relative coordinates carry at most e0 = 2^-24 ext of rounding, a
      // difference at most e = 2 e0 + 2^-24 1.01 rc, and the squared distance
      // near rc^2 at most 2 sqrt(D) rc e + D e^2 + (2.2 D) 2^-24 rc^2 (the
      // rounding of the D squares and their sum); the band is twice that.
      const float ext = fmaxf(fmaxf(s_ext[0], s_ext[1]), fmaxf(s_ext[2], s_ext[3]));
      const double rc = sqrt(rc2);
      const double e = 2.0 * ldexp((double)ext, -24) + ldexp(1.01 * rc, -24);
      const double band = 2.0 * (2.0 * sqrt((double)D) * rc * e + D * e * e + 2.2 * D * ldexp(rc2, -24));
      const float lo2 = (float)(rc2 - band) * (1.0f - 1.2e-7f), hi2 = (float)(rc2 + band) * (1.0f + 1.2e-7f);
      float xf[D];
#pragma unroll
      for (int d = 0; d < D; ++d) xf[d] = (float)(xi[d] - pos[tile * kHaloRows * D + d]);
      for (int o = 0; o < NOFF; ++o) {
        const int u = cand_u[hidx * NOFF + o];
        if (u < 0) continue;
#pragma unroll
        for (int wl = 0; wl < 2; ++wl) {
          const int c0 = wl ? uws[u] : ufs[u], len = wl ? uwl[u] : ufl[u], o0 = wl ? F + uwo[u] : ufo[u];
          for (int q0 = 0; q0 < len; q0 += 4) {
            HP cq[4];
#pragma unroll
            for (int uu = 0; uu < 4; ++uu) cq[uu] = hx[o0 + (q0 + uu < len ? q0 + uu : len - 1)];
#pragma unroll
            for (int uu = 0; uu < 4; ++uu) {
              const int q = q0 + uu;
              const int32_t j = c0 + q;
              if (q >= len || j == (int32_t)r) continue;
              float d2;
              if constexpr (D == 2) {
                const float dx = xf[0] - cq[uu].x, dy = xf[1] - cq[uu].y;
                d2 = dx * dx + dy * dy;
              } else {
                const float dx = xf[0] - cq[uu].x, dy = xf[1] - cq[uu].y, dz = xf[2] - cq[uu].z;
                d2 = dx * dx + dy * dy + dz * dz;
              }
              bool in = d2 < lo2;
              if (!in && d2 <= hi2) in = dist2_pt<double, D, false>(xi, pos + (int64_t)j * D, nullptr) <= rc2;
              if (in) accept(j, o0 + q);
            }
          }
        }
      }
    } else {
      cg.ranges<D>(r, true, [&](int32_t rb, int32_t re) {
        for (int32_t j = rb; j < re; ++j) {
          if (j == (int32_t)r) continue;
          if (dist2_pt<double, D, false>(xi, pos + (int64_t)j * D, nullptr) <= rc2) accept(j, 0);
        }
      });
    }
    if (FILL && (k & 3) && (k >> 2) < chunks) {
      const int sl = k & 3;                       // the row itself (halo: position 0) in the unused slots
      gb.y = sl > 1 ? gb.y : (int)r; gb.z = sl > 2 ? gb.z : (int)r; gb.w = (int)r;
      hb.y = sl > 1 ? hb.y : 0; hb.z = sl > 2 ? hb.z : 0; hb.w = 0;
      gp[(k >> 2) * 64] = gb;
      if (ok) hp[(k >> 2) * 64] = hb;
    }
    cnt[r] = k;
  }
  int m = k;
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(m, o, 64);
    m = v > m ? v : m;
  }
  if ((t & 63) == 0) s_kmax[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kHaloRows / 64; ++w) m = s_kmax[w] > m ? s_kmax[w] : m;
    blk[2 * tile] = (unsigned int)m;
    blk[2 * tile + 1] = ok ? (unsigned int)(F + W + 1) : 0u;
    halo_cnt[tile] = ok ? make_int2(F + W, F) : make_int2(-1, -1);
  }
}

// blk (ntiles pairs) and the wall rows' longest list (wall_kmax) -> *out; a
// list that did not fit its K slots, or a row outside the grid, marks both
// stale flags so that the steps queued on this build change nothing and the
// host builds again.  One block of kHaloRows threads.
static __global__ void __launch_bounds__(kHaloRows) k_halo_lists_reduce(
    const unsigned int* __restrict__ blk, int64_t ntiles, const unsigned int* __restrict__ wall_kmax,
    const unsigned int* __restrict__ outside, int64_t K, unsigned long long* __restrict__ stale,
    BuildStats* __restrict__ out) {
  unsigned int km = 0, hm = 0, used = 0;
  for (int64_t b = threadIdx.x; b < ntiles; b += blockDim.x) {
    const unsigned int a = blk[2 * b], h = blk[2 * b + 1];
    km = a > km ? a : km;
    hm = h > hm ? h : hm;
    used += h ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned int a = __shfl_xor(km, o, 64), h = __shfl_xor(hm, o, 64);
    km = a > km ? a : km;
    hm = h > hm ? h : hm;
    used += __shfl_xor(used, o, 64);
  }
  __shared__ unsigned int part[3][kHaloRows / 64];
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = km; part[1][threadIdx.x >> 6] = hm; part[2][threadIdx.x >> 6] = used;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kHaloRows / 64; ++w) {
      km = part[0][w] > km ? part[0][w] : km;
      hm = part[1][w] > hm ? part[1][w] : hm;
      used += part[2][w];
    }
    const unsigned int wk = *wall_kmax;
    km = wk > km ? wk : km;
    const unsigned int outv = *outside;
    out->kmax = km;
    out->hmax = hm ? hm - 1 : 0;
    out->used = used;
    out->outside = outv;
    out->K = (unsigned int)K;
    if ((int64_t)km > K || outv) { stale[0] = 1ull; stale[1] = 1ull; }
    __threadfence_system();
  }
}

}  // namespace tsph
