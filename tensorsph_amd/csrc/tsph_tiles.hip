// LDS cell-tile force sweep of the resident 2-D stepper (the density and
// wall sweeps of the same steps run over the per-row slot lists, which the
// rebuild here also writes).
//
// The neighbour search's cells (edge 2h + skin) are grouped into TX x TY
// tiles.  Rows are ordered tile-major (fluid rows first, then wall rows, each
// by tile then cell), and a dense table gives every cell's fluid and wall row
// ranges.  One workgroup takes one tile: it copies the (TX+2) x (TY+2) block
// of cells around the tile (the tile plus a one-cell halo) from HBM into LDS
// with contiguous, coalesced range reads, laid out so that the three cells
// of any stencil row are adjacent.  Each target particle of the tile then
// scans three contiguous LDS ranges, so the force sweep streams only the
// particle state itself.
//
// Visiting order is fixed (stencil rows bottom to top, cells left to right,
// fluid members then wall members, each in row order), so sums are
// reproducible and identical whether a tile runs from LDS or, when its
// block holds more particles than the LDS budget, straight from HBM.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "tsph_common.hpp"
#include "tsph_engine.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"

namespace tsph {

struct TileGeom {
  int64_t ext0, ext1;   // cells along x and y
  int64_t nty;          // tiles along y
  int TX, TY;           // cells per tile
  int64_t ncells;       // ext0 * ext1
};

constexpr int kMaxC = 18 * 6;   // up to 16 x 4 tiles: (16 + 2) x (4 + 2) block cells
// Per-thread list of accepted LDS slots (two-phase stencil scan): phase 1
// tests every stencil candidate's distance and keeps those within the
// support, phase 2 runs the pair physics over the kept slots only, so lanes
// of a wave do not all pay the physics on every rejected candidate.
constexpr int kNbMax = 32;

// ------------------------------------------------------------------ rebuild kernels
__global__ void k_tile_keys(const int64_t* __restrict__ cells, int64_t nt, int64_t n, int TX, int TY,
                            int64_t nty, unsigned long long K2, unsigned long long* __restrict__ keys,
                            int32_t* __restrict__ idx) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nt) return;
  const int64_t cx = cells[2 * r], cy = cells[2 * r + 1];
  const int64_t tile = (cx / TX) * nty + (cy / TY);
  const int64_t local = (cy % TY) * TX + (cx % TX);
  keys[r] = (r >= n ? K2 : 0ull) + (unsigned long long)(tile * (TX * TY) + local);
  idx[r] = (int32_t)r;
}

template <typename S>
__global__ void k_tile_permute(const int32_t* __restrict__ new2old, int64_t nt, const double* __restrict__ pos,
                               const double* __restrict__ vel, const S* __restrict__ mass,
                               const int64_t* __restrict__ orig, const int64_t* __restrict__ cells_old,
                               double* __restrict__ pos2, double* __restrict__ vel2, S* __restrict__ mass2,
                               int64_t* __restrict__ orig2, int64_t* __restrict__ cells_new, int64_t lo_off) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const int64_t o = new2old[k];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    pos2[k * 2 + d] = pos[o * 2 + d];
    vel2[k * 2 + d] = vel[o * 2 + d];
    cells_new[k * 2 + d] = cells_old[o * 2 + d];
  }
  mass2[k] = mass[o];
  if (lo_off) mass2[lo_off + k] = mass[lo_off + o];    // the low parts (FP32 runs)
  orig2[k] = orig[o];
}

__global__ void k_cell_ranges(const int64_t* __restrict__ cells, int64_t nt, int64_t n, int64_t ext1,
                              int64_t ncells, int32_t* __restrict__ R) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nt) return;
  const int64_t L = cells[2 * r] * ext1 + cells[2 * r + 1];
  const bool wall = r >= n;
  bool prev_same = false, next_same = false;
  if (r > 0 && ((r - 1) >= n) == wall) prev_same = cells[2 * (r - 1)] * ext1 + cells[2 * (r - 1) + 1] == L;
  if (r + 1 < nt && ((r + 1) >= n) == wall) next_same = cells[2 * (r + 1)] * ext1 + cells[2 * (r + 1) + 1] == L;
  if (!prev_same) R[(wall ? 2 : 0) * ncells + L] = (int32_t)r;
  if (!next_same) R[(wall ? 3 : 1) * ncells + L] = (int32_t)(r + 1);
}

TSPH_D int64_t tile_of(const unsigned long long* skeys, int64_t r, int64_t n, unsigned long long K2, int TT) {
  const unsigned long long k = skeys[r] - (r >= n ? K2 : 0ull);
  return (int64_t)(k / (unsigned long long)TT);
}

__global__ void k_tile_flags(const unsigned long long* __restrict__ skeys, int64_t nt, int64_t n,
                             unsigned long long K2, int TT, unsigned int* __restrict__ flags) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nt) return;
  bool start = r == 0 || r == n;
  if (!start) start = tile_of(skeys, r - 1, n, K2, TT) != tile_of(skeys, r, n, K2, TT);
  flags[r] = start ? 1u : 0u;
}

__global__ void k_tile_write(const unsigned long long* __restrict__ skeys, const unsigned int* __restrict__ flags,
                             const unsigned int* __restrict__ ids, int64_t nt, int64_t n, unsigned long long K2,
                             int TT, TileRec* __restrict__ rec) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nt) return;
  const int64_t id = (int64_t)ids[r] - 1;
  if (flags[r]) { rec[id].tile = tile_of(skeys, r, n, K2, TT); rec[id].start = (int32_t)r; }
  const bool last = r == nt - 1 || flags[r + 1];
  if (last) rec[id].end = (int32_t)(r + 1);
}

// ------------------------------------------------------------------ block staging
struct BlockTable {
  int off[kMaxC + 1];
  int fs[kMaxC], fn[kMaxC], ws[kMaxC];
};

// Fills the block's cell table (counts, then an exclusive prefix).  Returns
// the number of particles in the (TX+2) x (TY+2) block.
TSPH_D int stage_table(BlockTable& bt, const TileGeom& g, int64_t bx0, int64_t by0,
                       const int32_t* __restrict__ R) {
  const int BW = g.TX + 2, NC = BW * (g.TY + 2);
  for (int c = threadIdx.x; c < NC; c += blockDim.x) {
    const int64_t cx = bx0 + c % BW, cy = by0 + c / BW;
    int fs = 0, fe = 0, ws = 0, we = 0;
    if (cx >= 0 && cx < g.ext0 && cy >= 0 && cy < g.ext1) {
      const int64_t L = cx * g.ext1 + cy;
      fs = R[L]; fe = R[g.ncells + L]; ws = R[2 * g.ncells + L]; we = R[3 * g.ncells + L];
    }
    bt.fs[c] = fs; bt.fn[c] = fe - fs; bt.ws[c] = ws;
    bt.off[c + 1] = (fe - fs) + (we - ws);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    bt.off[0] = 0;
    for (int c = 0; c < NC; ++c) bt.off[c + 1] += bt.off[c];
  }
  __syncthreads();
  return bt.off[NC];
}

// row of the k-th particle of the staged block
TSPH_D int64_t staged_row(const BlockTable& bt, int NC, int k) {
  int lo = 0, hi = NC;                 // last c with off[c] <= k
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bt.off[mid] <= k) lo = mid; else hi = mid;
  }
  const int m = k - bt.off[lo];
  return m < bt.fn[lo] ? (int64_t)bt.fs[lo] + m : (int64_t)bt.ws[lo] + (m - bt.fn[lo]);
}

// Visit the 3 x 3 cell stencil of a target in block cell (bx, by): calls
// f(slot_or_row, from_lds).  Same order on both paths.
template <typename F>
TSPH_D void visit_stencil(const BlockTable& bt, int BW, int bx, int by, bool lds, F&& f) {
#pragma unroll 1
  for (int dy = -1; dy <= 1; ++dy) {
    const int c0 = (by + dy) * BW + (bx - 1);
    if (lds) {
      const int k1 = bt.off[c0 + 3];
      for (int k = bt.off[c0]; k < k1; ++k) f((int64_t)k);
    } else {
#pragma unroll 1
      for (int c = c0; c < c0 + 3; ++c) {
        const int nfl = bt.fn[c], nwl = bt.off[c + 1] - bt.off[c] - bt.fn[c];
        for (int m = 0; m < nfl; ++m) f((int64_t)bt.fs[c] + m);
        for (int m = 0; m < nwl; ++m) f((int64_t)bt.ws[c] + m);
      }
    }
  }
}

// ------------------------------------------------------------------ sweep B (force + integrate)
template <typename T> TSPH_D T wmax(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T> TSPH_D T wmin(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
TSPH_D double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename S, int MAXP, bool METERS>
__global__ void __launch_bounds__(kBlock) k_tile_force(WcsphArgs<S> a, TileGeom g,
    const TileRec* __restrict__ tiles, int64_t ntiles, const int32_t* __restrict__ R,
    const int64_t* __restrict__ cells, S* __restrict__ acc_out, StepArgs st, SimFlags* fl, int slot,
    double half_skin, unsigned long long* overflow) {
  const int64_t b = xcd_block(blockIdx.x, gridDim.x);
  constexpr bool kSplit = sizeof(S) == 4;
  __shared__ BlockTable bt;
  __shared__ double2 sx[MAXP];
  __shared__ float2 sxf[MAXP];      // FP32 positions relative to the tile origin (scan prefilter)
  __shared__ double2 sv[MAXP];
  __shared__ S sm[MAXP], spr[MAXP], sprl[MAXP], svol[MAXP];
  __shared__ int srow[MAXP];
  __shared__ double red[MS_NSUM][kBlock / 64];
  __shared__ unsigned short nb[kNbMax * kBlock];
  if (fl->stale[slot]) {
    if (b < ntiles) {
      const TileRec tr = tiles[b];
      for (int64_t r = tr.start + threadIdx.x; r < tr.end; r += blockDim.x) {
        st.pos_out[2 * r] = a.pos[2 * r]; st.pos_out[2 * r + 1] = a.pos[2 * r + 1];
        st.vel_out[2 * r] = a.vel[2 * r]; st.vel_out[2 * r + 1] = a.vel[2 * r + 1];
      }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) fl->stale[slot ^ 1] = 1;
    return;
  }
  if (b >= ntiles) return;
  const TileRec tr = tiles[b];
  const int64_t bx0 = (tr.tile / g.nty) * g.TX - 1, by0 = (tr.tile % g.nty) * g.TY - 1;
  const int BW = g.TX + 2, NC = BW * (g.TY + 2);
  const int total = stage_table(bt, g, bx0, by0, R);
  const bool lds = total <= MAXP;
  const double ox = a.pos[2 * (int64_t)tr.start], oy = a.pos[2 * (int64_t)tr.start + 1];
  if (lds) {
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
      const int64_t row = staged_row(bt, NC, k);
      sx[k] = make_double2(a.pos[2 * row], a.pos[2 * row + 1]);
      sxf[k] = make_float2((float)(a.pos[2 * row] - ox), (float)(a.pos[2 * row + 1] - oy));
      sv[k] = make_double2(a.vel[2 * row], a.vel[2 * row + 1]);
      sm[k] = a.mass[row];
      spr[k] = a.pr2[row];
      sprl[k] = kSplit ? a.aux[row].w : S(0);
      svol[k] = a.vol[row];
      srow[k] = (int)row;
    }
  } else if (threadIdx.x == 0) {
    atomicAdd(overflow, 1ull);
  }
  __syncthreads();
  // per-thread partial meters over this thread's targets
  double ms[MS_NSUM] = {0, 0, 0, 0, 0, 0, 0};
  double mx_disp = 0.0, mx_spd = 0.0, mx_rho = -1e300, mx_p = -1e300, mn_p = 1e300, mn_rho = 1e300;
  double mx_acc = 0.0, mx_y = -1e300, mn_y = 1e300, mx_x = -1e300;
  unsigned long long hits = 0;
  for (int64_t r = tr.start + threadIdx.x; r < tr.end; r += blockDim.x) {
    const double xi0 = a.pos[2 * r], xi1 = a.pos[2 * r + 1];
    const double vi0 = a.vel[2 * r], vi1 = a.vel[2 * r + 1];
    const int bx = (int)(cells[2 * r] - bx0), by = (int)(cells[2 * r + 1] - by0);
    const S pri = a.pr2[r];
    const S prli = kSplit ? a.aux[r].w : S(0);
    Sum<S> ap0, ap1, av0, av1;
    auto body = [&](int64_t q) {
      int64_t j;
      double xj0, xj1, vj0, vj1;
      S mj, prj, prlj, volj;
      if (lds) {
        j = srow[q]; xj0 = sx[q].x; xj1 = sx[q].y; vj0 = sv[q].x; vj1 = sv[q].y;
        mj = sm[q]; prj = spr[q]; prlj = sprl[q]; volj = svol[q];
      } else {
        j = q; xj0 = a.pos[2 * q]; xj1 = a.pos[2 * q + 1]; vj0 = a.vel[2 * q]; vj1 = a.vel[2 * q + 1];
        mj = a.mass[q]; prj = a.pr2[q]; prlj = kSplit ? a.aux[q].w : S(0); volj = a.vol[q];
      }
      if (j == r) return;
      const S d0 = (S)(xi0 - xj0), d1 = (S)(xi1 - xj1);
      const S r2 = d0 * d0 + d1 * d1;
      if (r2 > a.skip2) return;
      const S gg = cubic_g_r2<S>(r2, a.k);
      // FP32: the contrast form of the resident force sweep (k_step_force)
      const S cp = kSplit ? -mj * ((prj - pri) + (prlj - prli)) : -mj * (pri + prj);
      ap0.add(cp * (d0 * gg));
      ap1.add(cp * (d1 * gg));
      const S blend = j < a.n ? S(1) : a.friction;
      if (blend != S(0)) {
        const S dg = d0 * (d0 * gg) + d1 * (d1 * gg);
        const S relu = -dg > S(0) ? -dg : S(0);
        const S cv = -(a.nu2 * blend) * (volj * quot<S>(relu, r2 + a.eta2));
        av0.add(cv * (S)(vi0 - vj0));
        av1.add(cv * (S)(vi1 - vj1));
      }
    };
    bool dense = false;
    if (lds) {
      // phase 1: FP32 tile-relative prefilter, a superset of the body's test
      // (its rounding is ~1e-6 of the support; the margin is 1e-4); the body
      // re-tests exactly and skips the target itself, so sums are unchanged.
      // Measured 4 % faster per step than the exact FP64-delta scan it
      // replaces, despite the LDS it adds.
      int cnt = 0;
      const float xf0 = (float)(xi0 - ox), xf1 = (float)(xi1 - oy);
      const float lim = (float)(a.skip2d * (1.0 + 1.0e-4));
      visit_stencil(bt, BW, bx, by, true, [&](int64_t q) {
        const float2 pq = sxf[q];
        const float d0 = xf0 - pq.x, d1 = xf1 - pq.y;
        if (d0 * d0 + d1 * d1 > lim) return;
        if (cnt < kNbMax) nb[cnt * kBlock + threadIdx.x] = (unsigned short)q;
        ++cnt;
      });
      if (cnt <= kNbMax) {
        for (int e = 0; e < cnt; ++e) body((int64_t)nb[e * kBlock + threadIdx.x]);
        dense = true;
      }
    }
    if (!dense) visit_stencil(bt, BW, bx, by, lds, body);
    S acc0, acc1;
    if (kSplit) {
      const double qi = (double)pri + (double)prli;
      acc0 = (S)((hi(ap0) + hi(av0)) - 2.0 * qi * (double)a.gsum[2 * r]);
      acc1 = (S)(((hi(ap1) + hi(av1)) - 2.0 * qi * (double)a.gsum[2 * r + 1]) - (double)a.gravity);
    } else {
      acc0 = ap0.get() + av0.get();
      acc1 = (ap1.get() + av1.get()) + (-a.gravity);
    }
    if (acc_out) { acc_out[2 * r] = acc0; acc_out[2 * r + 1] = acc1; }
    // symplectic Euler + clamp (droplet_impact.py:1745-1756, :1293-1312)
    const double dt = fl->step_dt;
    double vn0 = vi0 + dt * (double)acc0, vn1 = vi1 + dt * (double)acc1;
    double xn0 = xi0 + dt * vn0, xn1 = xi1 + dt * vn1;
    const int h0 = (xn0 < st.lo[0] - st.tol) || (xn0 > st.hi[0] + st.tol);
    const int h1 = (xn1 < st.lo[1] - st.tol) || (xn1 > st.hi[1] + st.tol);
    xn0 = xn0 < st.lo[0] ? st.lo[0] : (xn0 > st.hi[0] ? st.hi[0] : xn0);
    xn1 = xn1 < st.lo[1] ? st.lo[1] : (xn1 > st.hi[1] ? st.hi[1] : xn1);
    if (h0) vn0 = 0.0;
    if (h1) vn1 = 0.0;
    st.pos_out[2 * r] = xn0; st.pos_out[2 * r + 1] = xn1;
    st.vel_out[2 * r] = vn0; st.vel_out[2 * r + 1] = vn1;
    hits += (h0 | h1) ? 1ull : 0ull;
    const double e0 = xn0 - st.anchor[2 * r], e1 = xn1 - st.anchor[2 * r + 1];
    mx_disp = fmax(mx_disp, sqrt(e0 * e0 + e1 * e1));
    const double spd2 = vi0 * vi0 + vi1 * vi1;
    mx_spd = fmax(mx_spd, spd2);
    const double rhoi = (double)a.rho[r];
    mx_rho = fmax(mx_rho, rhoi);
    if (METERS) {
      // the FP32 run's mass as its two parts
      const double mi = (double)a.mass[r] + (a.mass_lo ? (double)a.mass_lo[r] : 0.0), pi_ = (double)a.p[r];
      ms[MS_MASS] += mi; ms[MS_MX] += mi * vi0; ms[MS_MY] += mi * vi1;
      ms[MS_KE] += 0.5 * (mi * spd2); ms[MS_RHO] += rhoi; ms[MS_PE] += mi * st.gravity * xi1;
      mn_rho = fmin(mn_rho, rhoi); mx_p = fmax(mx_p, pi_); mn_p = fmin(mn_p, pi_);
      mx_acc = fmax(mx_acc, sqrt((double)acc0 * (double)acc0 + (double)acc1 * (double)acc1));
      mx_y = fmax(mx_y, xi1); mn_y = fmin(mn_y, xi1); mx_x = fmax(mx_x, xi0);
    }
  }
  // block-level reductions (every thread of the block reaches this point)
  const double wdisp = wmax(mx_disp), wspd = sqrt(wmax(mx_spd)), wrho = wmax(mx_rho);
  unsigned long long whits = hits;
  for (int o = 32; o > 0; o >>= 1) whits += __shfl_xor(whits, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ double bmax[3][kBlock / 64];
  __shared__ unsigned long long bhits[kBlock / 64];
  if (lane == 0) { bmax[0][wave] = wdisp; bmax[1][wave] = wspd; bmax[2][wave] = wrho; bhits[wave] = whits; }
  __syncthreads();
  if (threadIdx.x == 0) {                 // one atomic per block and quantity
    double bd = bmax[0][0], bv = bmax[1][0], br = bmax[2][0];
    unsigned long long bh = bhits[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      bd = fmax(bd, bmax[0][w]); bv = fmax(bv, bmax[1][w]); br = fmax(br, bmax[2][w]); bh += bhits[w];
    }
    atomicMax(&fl->maxes[MM_DISP], ord_enc(bd));
    atomicMax(&fl->maxes[MM_PEAK_SPEED], ord_enc(bv));
    atomicMax(&fl->maxes[MM_PEAK_RHO], ord_enc(br));
    if (bh) atomicAdd(&fl->escaped, bh);
    if (bd > half_skin) fl->stale[slot ^ 1] = 1;
  }
  if (METERS) {
#pragma unroll
    for (int k = 0; k < MS_NSUM; ++k) {
      const double v = wsum(ms[k]);
      if (lane == 0) red[k][wave] = v;
    }
    // row extrema, ordered so that all are maxima: speed, rho, p, -p, -rho, |acc|, y, -y, x
    double ext[9] = {wspd, wrho, wmax(mx_p), -wmin(mn_p), -wmin(mn_rho), wmax(mx_acc), wmax(mx_y),
                     -wmin(mn_y), wmax(mx_x)};
    __shared__ double bext[9][kBlock / 64];
    if (lane == 0)
      for (int k = 0; k < 9; ++k) bext[k][wave] = ext[k];
    __syncthreads();
    if (threadIdx.x < MS_NSUM) {
      double v = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[threadIdx.x][w];
      st.block_sums[b * MS_NSUM + threadIdx.x] = v;
    }
    if (threadIdx.x == 0) {                 // one atomic per block and extremum
      for (int k = 0; k < 9; ++k)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) ext[k] = fmax(ext[k], bext[k][w]);
      atomicMax(&fl->maxes[MM_SPEED], ord_enc(ext[0]));
      atomicMax(&fl->maxes[MM_RHO_MAX], ord_enc(ext[1]));
      atomicMax(&fl->maxes[MM_P_MAX], ord_enc(ext[2]));
      atomicMin(&fl->mins[MN_P_MIN], ord_enc(-ext[3]));
      atomicMin(&fl->mins[MN_RHO_MIN], ord_enc(-ext[4]));
      atomicMax(&fl->maxes[MM_ACC], ord_enc(ext[5]));
      atomicMax(&fl->maxes[MM_Y_MAX], ord_enc(ext[6]));
      atomicMin(&fl->mins[MN_Y_MIN], ord_enc(-ext[7]));
      atomicMax(&fl->maxes[MM_X_MAX], ord_enc(ext[8]));
    }
  }
}

// ------------------------------------------------------------------ pair count
// Unordered pairs with separation <= sqrt(r2max): fluid-fluid pairs counted at
// the lower row, fluid-wall pairs once (from the fluid end).
__global__ void __launch_bounds__(kBlock) k_tile_count(const double* __restrict__ pos, int64_t n, TileGeom g,
    const TileRec* __restrict__ tiles, int64_t ntiles, const int32_t* __restrict__ R,
    const int64_t* __restrict__ cells, double r2max, unsigned long long* __restrict__ cnt) {
  const int64_t b = blockIdx.x;
  if (b >= ntiles) return;
  __shared__ BlockTable bt;
  const TileRec tr = tiles[b];
  const int64_t bx0 = (tr.tile / g.nty) * g.TX - 1, by0 = (tr.tile % g.nty) * g.TY - 1;
  const int BW = g.TX + 2;
  stage_table(bt, g, bx0, by0, R);
  unsigned long long c = 0;
  for (int64_t r = tr.start + threadIdx.x; r < tr.end; r += blockDim.x) {
    const int bx = (int)(cells[2 * r] - bx0), by = (int)(cells[2 * r + 1] - by0);
    visit_stencil(bt, BW, bx, by, false, [&](int64_t j) {
      if (j < n && j <= r) return;
      const double d0 = pos[2 * r] - pos[2 * j], d1 = pos[2 * r + 1] - pos[2 * j + 1];
      if (d0 * d0 + d1 * d1 <= r2max) ++c;
    });
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

int tile_count_pairs(tsph_sim* sim, double radius, unsigned long long* cnt, hipStream_t stream) {
  TileGeom g{sim->ext[0], sim->ext[1], sim->ntile[1], sim->tile_dim[0], sim->tile_dim[1],
             sim->ext[0] * sim->ext[1]};
  if (sim->ntiles_f > 0)
    k_tile_count<<<(unsigned)sim->ntiles_f, kBlock, 0, stream>>>(
        sim->pos[sim->cur].as<double>(), sim->n, g, sim->tiles_f.as<TileRec>(), sim->ntiles_f,
        sim->ranges.as<int32_t>(), sim->cells_new.as<int64_t>(), radius * radius, cnt);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// ------------------------------------------------------------------ host side
template <int D, typename S>
int sim_rebuild_tiles(tsph_sim* sim, hipStream_t stream) {
  sim->bgrid_ok = false;   // the tile order moves the wall rows: the next slot-list build bins every row
  static_assert(D == 2, "cell tiles are 2-D");
  tsph_ctx* ctx = sim->ctx;
  const int64_t n = sim->n, nt = sim->nt;
  const double rc = 2.0 * sim->prm.op.h + sim->prm.skin;
  const double* pos = sim->pos[sim->cur].as<double>();
  TSPH_TRY((run_search<double, 2>(ctx, pos, nt, rc, nullptr, stream)));
  CellSearch& cs = ctx->search;
  const int TX = sim->tile_dim[0], TY = sim->tile_dim[1];
  sim->ext[0] = cs.ext[0]; sim->ext[1] = cs.ext[1];
  sim->ntile[0] = (cs.ext[0] + TX - 1) / TX;
  sim->ntile[1] = (cs.ext[1] + TY - 1) / TY;
  const int64_t ncells = cs.ext[0] * cs.ext[1];
  TSPH_REQUIRE(ncells < (1ll << 31), TSPH_ERR_ARG, "cell grid too large for the tile tables");
  const unsigned long long K2 = (unsigned long long)(sim->ntile[0] * sim->ntile[1]) * (TX * TY);
  TSPH_TRY(sim->tkeys.ensure(8 * nt));
  TSPH_TRY(sim->tkeys_sorted.ensure(8 * nt));
  TSPH_TRY(sim->tidx.ensure(8 * nt));              // int32 in + int32 out
  TSPH_TRY(sim->tflags.ensure(8 * (nt + 1)));      // flags + ids
  TSPH_TRY(sim->ranges.ensure(sizeof(int32_t) * 4 * ncells));
  const unsigned g = grid_for(nt);
  int32_t* idx_in = sim->tidx.as<int32_t>();
  int32_t* idx_out = idx_in + nt;
  k_tile_keys<<<g, kBlock, 0, stream>>>(cs.cells.as<int64_t>(), nt, n, TX, TY, sim->ntile[1], K2,
                                        sim->tkeys.as<unsigned long long>(), idx_in);
  TSPH_CHECK_LAUNCH();
  int end_bit = 1;
  while (end_bit < 63 && (1ull << end_bit) < 2 * K2) ++end_bit;
  size_t bytes = 0;
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, sim->tkeys.as<unsigned long long>(),
      sim->tkeys_sorted.as<unsigned long long>(), idx_in, idx_out, (int)nt, 0, end_bit, stream));
  TSPH_TRY(cs.sort_tmp.ensure(bytes));
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(cs.sort_tmp.ptr, bytes, sim->tkeys.as<unsigned long long>(),
      sim->tkeys_sorted.as<unsigned long long>(), idx_in, idx_out, (int)nt, 0, end_bit, stream));
  const int other = sim->cur ^ 1;
  k_tile_permute<S><<<g, kBlock, 0, stream>>>(idx_out, nt, pos, sim->vel[sim->cur].as<double>(), sim->mass.as<S>(),
                                              sim->orig.as<int64_t>(), cs.cells.as<int64_t>(),
                                              sim->pos[other].as<double>(), sim->vel[other].as<double>(),
                                              sim->mass2.as<S>(), sim->orig2.as<int64_t>(),
                                              sim->cells_new.as<int64_t>(), sizeof(S) == 4 ? nt : 0);
  TSPH_CHECK_LAUNCH();
  std::swap(sim->mass, sim->mass2);
  std::swap(sim->orig, sim->orig2);
  sim->cur = other;
  TSPH_CHECK_HIP(hipMemsetAsync(sim->ranges.ptr, 0, sizeof(int32_t) * 4 * ncells, stream));
  k_cell_ranges<<<g, kBlock, 0, stream>>>(sim->cells_new.as<int64_t>(), nt, n, cs.ext[1], ncells,
                                          sim->ranges.as<int32_t>());
  // per-row lists for the density and wall sweeps, walked over the same cells
  TSPH_TRY((build_sim_slots<2>(sim, sim->pos[sim->cur].as<double>(), rc, cs.ext, cs.stride, stream)));
  unsigned int* flags = sim->tflags.as<unsigned int>();
  unsigned int* ids = flags + nt;
  k_tile_flags<<<g, kBlock, 0, stream>>>(sim->tkeys_sorted.as<unsigned long long>(), nt, n, K2, TX * TY, flags);
  TSPH_CHECK_LAUNCH();
  bytes = 0;
  TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, flags, ids, (int)nt, stream));
  TSPH_TRY(ctx->scan_tmp.ensure(bytes));
  TSPH_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(ctx->scan_tmp.ptr, bytes, flags, ids, (int)nt, stream));
  unsigned int cnt[2] = {0, 0};
  TSPH_CHECK_HIP(hipMemcpyAsync(&cnt[0], ids + n - 1, 4, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(&cnt[1], ids + nt - 1, 4, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  sim->ntiles_f = cnt[0];
  sim->ntiles_w = (int64_t)cnt[1] - cnt[0];
  TSPH_TRY(sim->tiles_f.ensure(sizeof(TileRec) * (cnt[1] + 1)));
  k_tile_write<<<g, kBlock, 0, stream>>>(sim->tkeys_sorted.as<unsigned long long>(), flags, ids, nt, n, K2,
                                         TX * TY, sim->tiles_f.as<TileRec>());
  TSPH_CHECK_LAUNCH();
  // both ping-pong buffers carry the static wall rows
  const double* npos = sim->pos[sim->cur].as<double>();
  if (nt > n) {
    TSPH_CHECK_HIP(hipMemcpyAsync(sim->pos[sim->cur ^ 1].as<double>() + n * 2, npos + n * 2,
                                  sizeof(double) * (nt - n) * 2, hipMemcpyDeviceToDevice, stream));
    TSPH_CHECK_HIP(hipMemcpyAsync(sim->vel[sim->cur ^ 1].as<double>() + n * 2,
                                  sim->vel[sim->cur].as<double>() + n * 2, sizeof(double) * (nt - n) * 2,
                                  hipMemcpyDeviceToDevice, stream));
  }
  TSPH_CHECK_HIP(hipMemcpyAsync(sim->anchor.ptr, npos, sizeof(double) * n * 2, hipMemcpyDeviceToDevice, stream));
  SimFlags* fl = sim->flags.as<SimFlags>();
  TSPH_CHECK_HIP(hipMemsetAsync(&fl->stale[0], 0, sizeof(unsigned long long) * 2, stream));
  // meter partials are indexed by tile block
  TSPH_TRY(sim->block_sums.ensure(8 * MS_NSUM * (int64_t)grid_xcd(sim->ntiles_f * kBlock)));
  sim->rebuilds++;
  sim->built = true;
  return TSPH_OK;
}

template <typename S, int MAXP>
int queue_tiles(tsph_sim* sim, int64_t k, bool meters, hipStream_t stream) {
  const int slot = (int)(k & 1);
  SimFlags* fl = sim->flags.as<SimFlags>();
  WcsphArgs<S> a = make_wcsph_args<S>(sim->prm.op, 2);
  a.pos = sim->pos[sim->cur].as<double>();
  a.vel = sim->vel[sim->cur].as<double>();
  a.mass = sim->mass.as<S>();
  a.mass_lo = sizeof(S) == 4 ? sim->mass.as<S>() + sim->nt : nullptr;
  a.rho = sim->rho.as<S>(); a.p = sim->p.as<S>(); a.pr2 = sim->pr2.as<S>(); a.vol = sim->vol.as<S>();
  a.aux = sizeof(S) == 4 ? sim->aux.as<Vec4<S>>() : nullptr;     // q_lo of every row
  a.gsum = sizeof(S) == 4 ? sim->gsum.as<S>() : nullptr;
  TileGeom g{sim->ext[0], sim->ext[1], sim->ntile[1], sim->tile_dim[0], sim->tile_dim[1],
             sim->ext[0] * sim->ext[1]};
  const TileRec* tf = sim->tiles_f.as<TileRec>();
  const TileRec* tw = tf + sim->ntiles_f;
  const int32_t* R = sim->ranges.as<int32_t>();
  const int64_t* cells = sim->cells_new.as<int64_t>();
  unsigned long long* ovf = &fl->lds_overflow;
  const unsigned gf = grid_xcd(sim->ntiles_f * kBlock);
  const int r = (int)(k % kRing);
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][0], stream));
  TSPH_TRY((queue_slot_density_wall<2, S>(sim, k, stream)));
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][2], stream));
  StepArgs st{};
  st.pos_out = sim->pos[sim->cur ^ 1].as<double>();
  st.vel_out = sim->vel[sim->cur ^ 1].as<double>();
  for (int d = 0; d < 3; ++d) {
    st.lo[d] = sim->prm.clamp ? sim->prm.lo[d] : -1e300;
    st.hi[d] = sim->prm.clamp ? sim->prm.hi[d] : 1e300;
  }
  st.tol = sim->prm.tol;
  st.anchor = sim->anchor.as<double>();
  st.block_sums = sim->block_sums.as<double>();
  st.gravity = sim->prm.op.gravity;
  if (meters)
    k_tile_force<S, MAXP, true><<<gf, kBlock, 0, stream>>>(a, g, tf, sim->ntiles_f, R, cells, sim->acc.as<S>(),
                                                           st, fl, slot, 0.5 * sim->prm.skin, ovf);
  else
    k_tile_force<S, MAXP, false><<<gf, kBlock, 0, stream>>>(a, g, tf, sim->ntiles_f, R, cells, sim->acc.as<S>(),
                                                            st, fl, slot, 0.5 * sim->prm.skin, ovf);
  TSPH_CHECK_LAUNCH();
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][3], stream));
  return TSPH_OK;
}

template <int D, typename S>
int sim_queue_step_tiles(tsph_sim* sim, int64_t k, bool meters, hipStream_t stream) {
  static_assert(D == 2, "cell tiles are 2-D");
  const int cells = (sim->tile_dim[0] + 2) * (sim->tile_dim[1] + 2);
  // LDS budget by block size: ~8.6 particles per cell at h = 1.3 dx and cells
  // of 2.25 h, with headroom for compression; a block beyond it runs from HBM
  if (cells <= 36) return queue_tiles<S, 512>(sim, k, meters, stream);
  if (sizeof(S) == 8) return queue_tiles<S, 640>(sim, k, meters, stream);
  return queue_tiles<S, 768>(sim, k, meters, stream);
}

template int sim_rebuild_tiles<2, double>(tsph_sim*, hipStream_t);
template int sim_rebuild_tiles<2, float>(tsph_sim*, hipStream_t);
template int sim_queue_step_tiles<2, double>(tsph_sim*, int64_t, bool, hipStream_t);
template int sim_queue_step_tiles<2, float>(tsph_sim*, int64_t, bool, hipStream_t);

}  // namespace tsph
