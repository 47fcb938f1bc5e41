// Batch engine for served δ⁺-SPH simulations: many simulations of uneven size
// resident in HBM, advanced together by one captured step graph.
//
// Layout.  Fluid rows and wall stations of every resident simulation live in
// packed capacity-sized arrays; each simulation owns one contiguous fluid range
// and one contiguous wall range, and rows outside any simulation carry sim id
// -1.  Every kernel is launched over the capacity, so the step graph does not
// depend on which simulations are resident: admitting or releasing one only
// writes rows and table entries between graph replays.
//
// One step (all on the device, no host synchronisation):
//   1. per-simulation bounding box of cell coordinates (cell edge 2h)
//   2. segmented grid layout: per-simulation extents and first global cell
//   3. counting sort of all rows into cells, members sorted by row (stable)
//   4. neighbour lists (count, scan, fill): fluid targets see fluid and wall
//      rows of their own simulation, wall targets see fluid rows
//   5. the four δ⁺ sweeps (tsph_deltaplus.hip) over every simulation at once
//   6. per-simulation max |v|, step counter, adaptive dt for the next step
//      (courant 2 dx / (c0 + max|v|), benchmarks/delta_plus/elliptical_drop.py:50),
//      done / failed state, and a record copied to pinned host memory.
// A simulation's neighbour order depends only on its own rows, so its results
// are bitwise the same alone and in any mix.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tsph_common.hpp"
#include "tsph_deltaplus.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"

namespace tsph {

constexpr int kSegSlots = 8;       // distinct simulations reduced in LDS per block
constexpr int kScanTile = 1024;
constexpr unsigned long long kCellBias = 1ull << 62;

enum SimState : unsigned char { SS_FREE = 0, SS_ACTIVE = 1, SS_DONE = 2, SS_FAILED = 3 };

struct BSim {                       // device-side bookkeeping of one simulation
  double cell;                      // cell edge = list radius = 2h
  double courant;
  int64_t budget, steps, done_step;
  int64_t cmin[2], ext[2];
  int64_t base;                     // first global cell id of this simulation's grid
};

struct BatchDev {
  int64_t cap_f, cap_w, cap_s, cap_cells;
  int64_t kf, kw;                   // list slots per fluid row / per wall station
  const double *x, *v, *rho;        // current state (ping-pong side)
  double *x_new, *v_new, *rho_new;
  const double *m, *xw, *vw, *nw_;
  const int32_t *fsim, *wsim;
  DpSim* dps;
  BSim* bs;
  unsigned char* state;
  const unsigned char* active;      // 1 while SS_ACTIVE (kept beside state for the sweeps)
  int* status;
  unsigned long long *vmax, *bbox;  // bbox: 4 per simulation (min x, min y, max x, max y), biased
  int64_t* total_cells;
  unsigned int* flags;              // bit 0 cell overflow
  int32_t *cid, *ccount, *cfill, *members;
  int64_t* cstart;
  // Lists in fixed slots, interleaved by wave in chunks of four: entry k of
  // row r sits at int32 index (r / 64) * 64 * K + 256 (k / 4) + 4 (r % 64)
  // + k % 4 (K a multiple of 4), so a lane writes its entries four at a time
  // as one 16-byte store, the 64 rows of a wave read their entries 4c .. 4c+3
  // from one 1 KB run, and no count pass or scan is needed; fcnt / wcnt hold
  // the counts.
  int32_t *fcnt, *wcnt, *fidx, *widx;
  int64_t* step_counter;
};

TSPH_D bool row_live(const BatchDev& b, int64_t c, int& s) {
  // c: cloud index, fluid rows [0, cap_f), wall stations cap_f + w
  s = c < b.cap_f ? b.fsim[c] : b.wsim[c - b.cap_f];
  return s >= 0 && b.active[s];
}

TSPH_D void row_pos(const BatchDev& b, int64_t c, double& x, double& y) {
  if (c < b.cap_f) { x = b.x[2 * c]; y = b.x[2 * c + 1]; }
  else { const int64_t w = c - b.cap_f; x = b.xw[2 * w]; y = b.xw[2 * w + 1]; }
}

TSPH_D int64_t cell_of(double x, double cell) { return (int64_t)floor(x / cell); }

// ---------------------------------------------------------------- 1. bounding boxes
__global__ void k_b_bbox_reset(BatchDev b) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.cap_s) return;
  b.bbox[4 * s] = ~0ull; b.bbox[4 * s + 1] = ~0ull;
  b.bbox[4 * s + 2] = 0ull; b.bbox[4 * s + 3] = 0ull;
  if (s == 0) *b.flags = 0u;
}

// Rows of one simulation are contiguous, so a block of rows touches few
// simulations: reduce in LDS per simulation, then one global atomic each.
__global__ void __launch_bounds__(kBlock) k_b_bbox(BatchDev b) {
  __shared__ unsigned long long sm[kSegSlots][4];
  __shared__ int s0;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncloud = b.cap_f + b.cap_w;
  if (threadIdx.x < kSegSlots) {
    sm[threadIdx.x][0] = ~0ull; sm[threadIdx.x][1] = ~0ull; sm[threadIdx.x][2] = 0ull; sm[threadIdx.x][3] = 0ull;
  }
  if (threadIdx.x == 0) {
    const int64_t c0 = (int64_t)blockIdx.x * blockDim.x;
    int s = -1;
    if (c0 < ncloud) s = c0 < b.cap_f ? b.fsim[c0] : b.wsim[c0 - b.cap_f];
    s0 = s < 0 ? 0 : s;
  }
  __syncthreads();
  int s;
  if (c < ncloud && row_live(b, c, s)) {
    double x, y;
    row_pos(b, c, x, y);
    const double cell = b.bs[s].cell;
    const unsigned long long ex = (unsigned long long)(cell_of(x, cell) + (int64_t)kCellBias);
    const unsigned long long ey = (unsigned long long)(cell_of(y, cell) + (int64_t)kCellBias);
    const int k = s - s0;
    if (k >= 0 && k < kSegSlots) {
      atomicMin(&sm[k][0], ex); atomicMin(&sm[k][1], ey);
      atomicMax(&sm[k][2], ex); atomicMax(&sm[k][3], ey);
    } else {
      atomicMin(&b.bbox[4 * s], ex); atomicMin(&b.bbox[4 * s + 1], ey);
      atomicMax(&b.bbox[4 * s + 2], ex); atomicMax(&b.bbox[4 * s + 3], ey);
    }
  }
  __syncthreads();
  if (threadIdx.x < kSegSlots && sm[threadIdx.x][2] != 0ull) {
    const int64_t t = s0 + threadIdx.x;
    atomicMin(&b.bbox[4 * t], sm[threadIdx.x][0]); atomicMin(&b.bbox[4 * t + 1], sm[threadIdx.x][1]);
    atomicMax(&b.bbox[4 * t + 2], sm[threadIdx.x][2]); atomicMax(&b.bbox[4 * t + 3], sm[threadIdx.x][3]);
  }
}

// ---------------------------------------------------------------- 2. grid layout (one block)
__global__ void __launch_bounds__(1024) k_b_layout(BatchDev b) {
  __shared__ int64_t part[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t s0 = 0; s0 < b.cap_s; s0 += blockDim.x) {
    const int64_t s = s0 + threadIdx.x;
    int64_t nc = 0;
    if (s < b.cap_s && b.active[s] && b.bbox[4 * s + 2] != 0ull) {
      BSim& q = b.bs[s];
      for (int d = 0; d < 2; ++d) {
        q.cmin[d] = (int64_t)(b.bbox[4 * s + d] - kCellBias);
        q.ext[d] = (int64_t)(b.bbox[4 * s + 2 + d] - kCellBias) - q.cmin[d] + 1;
      }
      nc = q.ext[0] * q.ext[1];
    }
    part[threadIdx.x] = nc;
    __syncthreads();
    for (int st = 1; st < (int)blockDim.x; st <<= 1) {       // inclusive scan
      const int64_t v = threadIdx.x >= (unsigned)st ? part[threadIdx.x - st] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    if (s < b.cap_s) b.bs[s].base = carry + part[threadIdx.x] - nc;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += part[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *b.total_cells = carry;
    if (carry > b.cap_cells) atomicOr(b.flags, 1u);
  }
}

// ---------------------------------------------------------------- 3. counting sort into cells
__global__ void k_b_zero_cells(BatchDev b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.cap_cells) return;
  b.ccount[k] = 0;
  b.cfill[k] = 0;
}

__global__ void __launch_bounds__(kBlock) k_b_cell_count(BatchDev b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.cap_f + b.cap_w) return;
  int s;
  if (!row_live(b, c, s) || (*b.flags & 1u)) { b.cid[c] = -1; return; }
  double x, y;
  row_pos(b, c, x, y);
  const BSim& q = b.bs[s];
  const int64_t cx = cell_of(x, q.cell) - q.cmin[0], cy = cell_of(y, q.cell) - q.cmin[1];
  const int64_t id = q.base + cx * q.ext[1] + cy;
  b.cid[c] = (int32_t)id;
  atomicAdd(&b.ccount[id], 1);
}

// exclusive scan of n_dev int32 counts (n_dev read on the device, at most cap)
// into int64 offsets out[0..n], over a grid sized by cap (graph-stable)
__global__ void __launch_bounds__(kScanTile) k_scan_tiles(const int32_t* __restrict__ cnt, const int64_t* n_dev,
                                                          int64_t n_fixed, int64_t* __restrict__ out,
                                                          int64_t* __restrict__ tile_sums) {
  __shared__ int64_t sh[kScanTile];
  int64_t n = n_dev ? *n_dev : n_fixed;
  n = n < n_fixed ? n : n_fixed;                 // never past the capacity
  const int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
  const int64_t v = i < n ? (int64_t)cnt[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int st = 1; st < kScanTile; st <<= 1) {
    const int64_t u = threadIdx.x >= (unsigned)st ? sh[threadIdx.x - st] : 0;
    __syncthreads();
    sh[threadIdx.x] += u;
    __syncthreads();
  }
  out[i] = sh[threadIdx.x] - v;                  // tile-local exclusive
  if (threadIdx.x == kScanTile - 1) tile_sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void __launch_bounds__(1024) k_scan_sums(int64_t* __restrict__ tile_sums, int64_t ntiles) {
  __shared__ int64_t part[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t t0 = 0; t0 < ntiles; t0 += blockDim.x) {
    const int64_t t = t0 + threadIdx.x;
    const int64_t v = t < ntiles ? tile_sums[t] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int st = 1; st < (int)blockDim.x; st <<= 1) {
      const int64_t u = threadIdx.x >= (unsigned)st ? part[threadIdx.x - st] : 0;
      __syncthreads();
      part[threadIdx.x] += u;
      __syncthreads();
    }
    if (t < ntiles) tile_sums[t] = carry + part[threadIdx.x] - v;   // exclusive tile offset
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += part[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

__global__ void k_scan_add(int64_t* __restrict__ out, const int64_t* __restrict__ tile_sums, const int64_t* n_dev,
                           int64_t n_fixed, int64_t ntiles) {
  int64_t n = n_dev ? *n_dev : n_fixed;
  n = n < n_fixed ? n : n_fixed;
  const int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
  if (i < n) out[i] += tile_sums[blockIdx.x];
  if (i == 0) out[n] = tile_sums[ntiles];        // total after the last element
}

struct ScanPlan {
  int64_t cap, ntiles;
  int64_t* tiles;                                // ntiles + 1
};

int queue_scan(const ScanPlan& p, const int32_t* cnt, const int64_t* n_dev, int64_t n_fixed, int64_t* out,
               hipStream_t stream) {
  k_scan_tiles<<<(unsigned)p.ntiles, kScanTile, 0, stream>>>(cnt, n_dev, n_fixed, out, p.tiles);
  k_scan_sums<<<1, 1024, 0, stream>>>(p.tiles, p.ntiles);
  k_scan_add<<<(unsigned)p.ntiles, kScanTile, 0, stream>>>(out, p.tiles, n_dev, n_fixed, p.ntiles);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

__global__ void __launch_bounds__(kBlock) k_b_scatter(BatchDev b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.cap_f + b.cap_w) return;
  const int32_t id = b.cid[c];
  if (id < 0) return;
  const int64_t slot = b.cstart[id] + atomicAdd(&b.cfill[id], 1);
  b.members[slot] = (int32_t)c;
}

// members of each cell in ascending row order (the scatter order is arbitrary)
__global__ void __launch_bounds__(kBlock) k_b_cell_sort(BatchDev b) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= b.cap_cells || id >= *b.total_cells || (*b.flags & 1u)) return;
  int32_t* seg = b.members + b.cstart[id];
  const int n = b.ccount[id];
  for (int a = 1; a < n; ++a) {
    const int32_t v = seg[a];
    int k = a - 1;
    while (k >= 0 && seg[k] > v) { seg[k + 1] = seg[k]; --k; }
    seg[k + 1] = v;
  }
}

// ---------------------------------------------------------------- 4. neighbour lists
// MODE 0 counts, MODE 1 writes.  Fluid rows: every other live row of the same
// simulation within the list radius (fluid rows sort before walls in a cell);
// wall stations: fluid rows only.  Visit order: 3 x 3 cells lexicographic.
// One pass over the 3 x 3 cells around every row: fluid rows list fluid and
// wall neighbours of their own simulation, wall stations list fluid rows, all
// with d^2 <= (cell edge)^2, into the row's fixed slots.  A row with more
// neighbours than slots keeps the first ones and fails its simulation
// (status bit 2): its step is discarded by k_b_advance.
__global__ void __launch_bounds__(kBlock) k_b_lists(BatchDev b) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= b.cap_f + b.cap_w) return;
  const bool wall = c >= b.cap_f;
  const int64_t r = wall ? c - b.cap_f : c;
  int32_t* cnt = wall ? b.wcnt : b.fcnt;
  if (b.cid[c] < 0) { cnt[r] = 0; return; }
  int s;
  row_live(b, c, s);
  double x, y;
  row_pos(b, c, x, y);
  const BSim& q = b.bs[s];
  const double cut2 = q.cell * q.cell;
  const int64_t cx = cell_of(x, q.cell) - q.cmin[0], cy = cell_of(y, q.cell) - q.cmin[1];
  const int64_t K = wall ? b.kw : b.kf;
  int4* out = reinterpret_cast<int4*>(wall ? b.widx : b.fidx) + (r >> 6) * 16 * K + (r & 63);
  const int64_t chunks = K >> 2;
  int4 buf = make_int4((int)c, (int)c, (int)c, (int)c);
  int64_t k = 0;
  for (int dx = -1; dx <= 1; ++dx) {
    const int64_t nx = cx + dx;
    if (nx < 0 || nx >= q.ext[0]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t ny = cy + dy;
      if (ny < 0 || ny >= q.ext[1]) continue;
      const int64_t id = q.base + nx * q.ext[1] + ny;
      const int64_t m0 = b.cstart[id], m1 = m0 + b.ccount[id];
      // candidates four at a time: member ids, then their positions, then the tests
      for (int64_t e0 = m0; e0 < m1; e0 += 4) {
        int64_t o4[4];
        double xo[4], yo[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) o4[u] = b.members[e0 + u < m1 ? e0 + u : m1 - 1];   // m1 > e0 >= m0
#pragma unroll
        for (int u = 0; u < 4; ++u) row_pos(b, o4[u], xo[u], yo[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t o = o4[u];
          if (e0 + u >= m1 || o == c) continue;
          if (wall && o >= b.cap_f) continue;
          double d2;
          {
#pragma clang fp contract(off)
            const double sx = x - xo[u], sy = y - yo[u];
            d2 = sx * sx + sy * sy;
          }
          if (d2 > cut2) continue;
          const int sl = (int)(k & 3);
          buf.x = sl == 0 ? (int)o : buf.x;
          buf.y = sl == 1 ? (int)o : buf.y;
          buf.z = sl == 2 ? (int)o : buf.z;
          buf.w = sl == 3 ? (int)o : buf.w;
          if (sl == 3 && (k >> 2) < chunks) out[(k >> 2) * 64] = buf;
          ++k;
        }
      }
    }
  }
  if ((k & 3) && (k >> 2) < chunks) {        // the last partial chunk; the row itself fills it up
    const int sl = (int)(k & 3);
    buf.y = sl > 1 ? buf.y : (int)c;
    buf.z = sl > 2 ? buf.z : (int)c;
    buf.w = (int)c;
    out[(k >> 2) * 64] = buf;
  }
  cnt[r] = (int32_t)(k < K ? k : K);
  if (k > K) atomicOr(&b.status[s], 4);
}

// Optional: each list in ascending row order (the order of tsph_nlist_from_pairs),
// so sums run in the single-step API's order; used to test the engine bitwise.
__global__ void __launch_bounds__(kBlock) k_b_sort_lists(BatchDev b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.cap_f + b.cap_w || b.cid[c] < 0) return;
  const bool wall = c >= b.cap_f;
  const int64_t r = wall ? c - b.cap_f : c;
  const int64_t K = wall ? b.kw : b.kf;
  int32_t* seg = (wall ? b.widx : b.fidx) + (r >> 6) * 64 * K + 4 * (r & 63);
  auto at = [&](int64_t k) -> int32_t& { return seg[256 * (k >> 2) + (k & 3)]; };   // chunked slots
  const int64_t n = (wall ? b.wcnt : b.fcnt)[r];
  for (int64_t a = 1; a < n; ++a) {
    const int32_t v = at(a);
    int64_t k = a - 1;
    while (k >= 0 && at(k) > v) { at(k + 1) = at(k); --k; }
    at(k + 1) = v;
  }
}

// ---------------------------------------------------------------- 6. per-simulation advance
__global__ void __launch_bounds__(kBlock) k_b_vmax(BatchDev b) {
  __shared__ unsigned long long sm[kSegSlots];
  __shared__ int s0;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (threadIdx.x < kSegSlots) sm[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) {
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x;
    const int s = t0 < b.cap_f ? b.fsim[t0] : -1;
    s0 = s < 0 ? 0 : s;
  }
  __syncthreads();
  if (t < b.cap_f) {
    const int s = b.fsim[t];
    if (s >= 0 && b.active[s]) {
#pragma clang fp contract(off)
      const double u = b.v_new[2 * t], w = b.v_new[2 * t + 1];
      const unsigned long long e = ord_enc(sqrt(u * u + w * w));
      const int k = s - s0;
      if (k >= 0 && k < kSegSlots) atomicMax(&sm[k], e);
      else atomicMax(&b.vmax[s], e);
    }
  }
  __syncthreads();
  if (threadIdx.x < kSegSlots && sm[threadIdx.x] != 0ull) atomicMax(&b.vmax[s0 + threadIdx.x], sm[threadIdx.x]);
}

TSPH_D double ord_dec(unsigned long long u) {
  const unsigned long long bits = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  return __longlong_as_double((long long)bits);
}

__global__ void k_b_advance(BatchDev b) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t step = *b.step_counter;
  if (s < b.cap_s && b.state[s] == SS_ACTIVE) {
    BSim& q = b.bs[s];
    DpSim& P = b.dps[s];
    q.steps += 1;
    if (b.status[s] != 0 || (*b.flags != 0u)) {
      b.state[s] = SS_FAILED;
      ((unsigned char*)b.active)[s] = 0;
      q.done_step = step;
    } else if (q.steps >= q.budget) {
      b.state[s] = SS_DONE;
      ((unsigned char*)b.active)[s] = 0;
      q.done_step = step;
    } else {
      const double vm = b.vmax[s] ? ord_dec(b.vmax[s]) : 0.0;
      P.dt = q.courant * 2.0 * P.dx / (P.c0 + vm);
    }
    b.vmax[s] = 0ull;
  }
}

// Record copied to the host after each step.
struct BatchRecord {
  int64_t step;
  unsigned int flags;
  unsigned int pad;
};

// after every per-simulation update of the step: count it and record it
__global__ void k_b_record(BatchDev b, BatchRecord* rec) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int64_t step = *b.step_counter + 1;
    *b.step_counter = step;
    rec->step = step;
    rec->flags = *b.flags;
  }
}

}  // namespace tsph

using namespace tsph;

// ====================================================================== host engine
struct Range { int64_t start, count; };

struct tsph_batch {
  tsph_batch_caps caps{};
  DevBuf x[2], v[2], rho[2], m, fsim, xw, vw, nw, wsim;
  DevBuf dps, bs, state, active, status, vmax, bbox, total_cells, flags, step_counter;
  DevBuf cid, ccount, cfill, members, cstart, cstart_tiles;
  DevBuf fcnt, wcnt, fidx, widx;
  DevBuf wall_p, wall_rho, reaction, acc, lam, inv, gr, normal, strength, shift, surface, vol;
  DevBuf record;
  BatchRecord* host_rec = nullptr;               // pinned
  unsigned char* host_state = nullptr;           // pinned, cap_s
  int64_t* host_steps = nullptr;                 // pinned, cap_s
  int cur = 0;
  int64_t steps_done = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipStream_t graph_stream = nullptr;
  // launch shape (tsph_batch_launch): rows per workgroup of the pair sweeps
  // and of the list build, XCD-chunked row blocks of the pair sweeps
  int blk_sweep = 256, blk_lists = 256, xcd = 1;
  std::vector<Range> fslot, wslot;               // per simulation slot
  std::vector<unsigned char> used;               // slot in use
  std::vector<int64_t> done_at;                  // host copy of the finishing step
  int64_t f_hw = 0, w_hw = 0;                    // high-water marks
};

namespace tsph {

static BatchDev batch_dev(tsph_batch* B, int side) {
  BatchDev d{};
  const tsph_batch_caps& c = B->caps;
  d.cap_f = c.fluid; d.cap_w = c.wall; d.cap_s = c.sims; d.cap_cells = c.cells;
  d.kf = (c.fluid_entries / c.fluid) & ~3ll;
  d.kw = c.wall > 0 ? (c.wall_entries / c.wall) & ~3ll : 0;
  d.x = B->x[side].as<double>(); d.v = B->v[side].as<double>(); d.rho = B->rho[side].as<double>();
  d.x_new = B->x[side ^ 1].as<double>(); d.v_new = B->v[side ^ 1].as<double>();
  d.rho_new = B->rho[side ^ 1].as<double>();
  d.m = B->m.as<double>(); d.xw = B->xw.as<double>(); d.vw = B->vw.as<double>(); d.nw_ = B->nw.as<double>();
  d.fsim = B->fsim.as<int32_t>(); d.wsim = B->wsim.as<int32_t>();
  d.dps = B->dps.as<DpSim>(); d.bs = B->bs.as<BSim>();
  d.state = B->state.as<unsigned char>(); d.active = B->active.as<unsigned char>();
  d.status = B->status.as<int>(); d.vmax = B->vmax.as<unsigned long long>();
  d.bbox = B->bbox.as<unsigned long long>(); d.total_cells = B->total_cells.as<int64_t>();
  d.flags = B->flags.as<unsigned int>();
  d.cid = B->cid.as<int32_t>(); d.ccount = B->ccount.as<int32_t>(); d.cfill = B->cfill.as<int32_t>();
  d.members = B->members.as<int32_t>(); d.cstart = B->cstart.as<int64_t>();
  d.fcnt = B->fcnt.as<int32_t>(); d.wcnt = B->wcnt.as<int32_t>();
  d.fidx = B->fidx.as<int32_t>(); d.widx = B->widx.as<int32_t>();
  d.step_counter = B->step_counter.as<int64_t>();
  return d;
}

static DpArgs batch_dp_args(tsph_batch* B, const BatchDev& d) {
  DpArgs a{};
  a.nf = d.cap_f; a.nw = d.cap_w;
  a.sims = d.dps; a.fsim = d.fsim; a.wsim = d.wsim; a.active = d.active; a.status = d.status;
  a.x = d.x; a.v = d.v; a.rho = d.rho; a.m = d.m;
  a.xw = d.xw; a.vw = d.vw; a.nw_ = d.nw_;
  a.wall_p = B->wall_p.as<double>(); a.wall_rho = B->wall_rho.as<double>();
  a.reaction = B->reaction.as<double>();
  a.acc = B->acc.as<double>(); a.vnew = d.v_new; a.lam = B->lam.as<double>(); a.inv = B->inv.as<double>();
  a.gr = B->gr.as<double>(); a.rho_new = d.rho_new; a.normal = B->normal.as<double>();
  a.strength = B->strength.as<double>(); a.shift = B->shift.as<double>(); a.x_new = d.x_new;
  a.surface = B->surface.as<unsigned char>();
  a.vol = B->vol.as<double>();
  a.fcnt = d.fcnt; a.wcnt = d.wcnt; a.kf = d.kf; a.kw = d.kw;
  return a;
}

static unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// queue one full step reading side `side`
static int batch_queue_step(tsph_batch* B, int side, hipStream_t stream) {
  const BatchDev d = batch_dev(B, side);
  const int64_t ncloud = d.cap_f + d.cap_w;
  k_b_bbox_reset<<<nblk(d.cap_s), kBlock, 0, stream>>>(d);
  k_b_bbox<<<nblk(ncloud), kBlock, 0, stream>>>(d);
  k_b_layout<<<1, 1024, 0, stream>>>(d);
  k_b_zero_cells<<<nblk(d.cap_cells), kBlock, 0, stream>>>(d);
  k_b_cell_count<<<nblk(ncloud), kBlock, 0, stream>>>(d);
  TSPH_CHECK_LAUNCH();
  ScanPlan pc{d.cap_cells, (d.cap_cells + kScanTile) / kScanTile, B->cstart_tiles.as<int64_t>()};
  TSPH_TRY(queue_scan(pc, d.ccount, d.total_cells, d.cap_cells, d.cstart, stream));
  k_b_scatter<<<nblk(ncloud), kBlock, 0, stream>>>(d);
  k_b_cell_sort<<<nblk(d.cap_cells), kBlock, 0, stream>>>(d);
  k_b_lists<<<grid_xcd(ncloud, B->blk_lists), B->blk_lists, 0, stream>>>(d);
  if (B->caps.sorted_lists) k_b_sort_lists<<<nblk(ncloud), kBlock, 0, stream>>>(d);
  TSPH_CHECK_LAUNCH();
  // the δ⁺ step lists index walls as n_fluid + w: here n_fluid = cap_f
  DpArgs a = batch_dp_args(B, d);
  a.block = B->blk_sweep;
  a.xcd = B->xcd;
  TSPH_TRY(dp_queue_step(a, nullptr, d.fidx, nullptr, d.widx, stream));
  k_b_vmax<<<nblk(d.cap_f), kBlock, 0, stream>>>(d);
  k_b_advance<<<nblk(d.cap_s), kBlock, 0, stream>>>(d);
  k_b_record<<<1, 64, 0, stream>>>(d, B->record.as<BatchRecord>());
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipMemcpyAsync(B->host_rec, B->record.ptr, sizeof(BatchRecord), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(B->host_state, B->state.ptr, d.cap_s, hipMemcpyDeviceToHost, stream));
  return TSPH_OK;
}

static int batch_capture(tsph_batch* B, hipStream_t stream) {
  for (int side = 0; side < 2; ++side) {
    if (B->graph[side]) { (void)hipGraphExecDestroy(B->graph[side]); B->graph[side] = nullptr; }
    hipGraph_t g = nullptr;
    TSPH_CHECK_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_queue_step(B, side, stream);
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc != TSPH_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    TSPH_CHECK_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&B->graph[side], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TSPH_CHECK_HIP(ei);
  }
  B->graph_stream = stream;
  return TSPH_OK;
}

// first-fit range in [0, cap) avoiding the ranges of used slots
static int64_t find_range(const std::vector<Range>& slots, const std::vector<unsigned char>& used, int64_t cap,
                          int64_t need) {
  if (need == 0) return 0;
  std::vector<Range> taken;
  for (size_t k = 0; k < slots.size(); ++k)
    if (used[k] && slots[k].count > 0) taken.push_back(slots[k]);
  std::sort(taken.begin(), taken.end(), [](const Range& a, const Range& b) { return a.start < b.start; });
  int64_t at = 0;
  for (const Range& r : taken) {
    if (r.start - at >= need) return at;
    at = std::max(at, r.start + r.count);
  }
  return cap - at >= need ? at : -1;
}

__global__ void k_b_fill_sim(int32_t* ids, int64_t start, int64_t count, int32_t value) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) ids[start + k] = value;
}

// max |v| over one simulation's rows (admission: the first step's dt)
__global__ void __launch_bounds__(1024) k_b_vmax_range(const double* v, int64_t start, int64_t count,
                                                       unsigned long long* out) {
  __shared__ double red[1024];
  double m = 0.0;
  for (int64_t k = threadIdx.x; k < count; k += blockDim.x) {
#pragma clang fp contract(off)
    const double u = v[2 * (start + k)], w = v[2 * (start + k) + 1];
    m = fmax(m, sqrt(u * u + w * w));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = ord_enc(red[0]);
}

}  // namespace tsph

extern "C" {

int tsph_batch_destroy(tsph_batch* B) {
  if (!B) return TSPH_OK;
  for (int s = 0; s < 2; ++s) if (B->graph[s]) (void)hipGraphExecDestroy(B->graph[s]);
  DevBuf* bufs[] = {&B->x[0], &B->x[1], &B->v[0], &B->v[1], &B->rho[0], &B->rho[1], &B->m, &B->fsim, &B->xw,
                    &B->vw, &B->nw, &B->wsim, &B->dps, &B->bs, &B->state, &B->active, &B->status, &B->vmax,
                    &B->bbox, &B->total_cells, &B->flags, &B->step_counter, &B->cid, &B->ccount, &B->cfill,
                    &B->members, &B->cstart, &B->cstart_tiles, &B->fcnt, &B->wcnt,
                    &B->fidx, &B->widx, &B->wall_p, &B->wall_rho, &B->reaction,
                    &B->acc, &B->lam, &B->inv, &B->gr, &B->normal, &B->strength, &B->shift, &B->surface,
                    &B->vol, &B->record};
  for (DevBuf* b : bufs) b->release();
  if (B->host_rec) (void)hipHostFree(B->host_rec);
  if (B->host_state) (void)hipHostFree(B->host_state);
  if (B->host_steps) (void)hipHostFree(B->host_steps);
  delete B;
  return TSPH_OK;
}

int tsph_batch_create(const tsph_batch_caps* caps, hipStream_t stream, tsph_batch** out) {
  TSPH_REQUIRE(caps && out, TSPH_ERR_ARG, "NULL argument");
  *out = nullptr;
  const tsph_batch_caps& c = *caps;
  TSPH_REQUIRE(c.fluid > 0 && c.wall >= 0 && c.sims > 0 && c.cells > 0 && c.fluid_entries > 0 &&
               c.wall_entries >= 0, TSPH_ERR_ARG, "capacities must be positive");
  TSPH_REQUIRE(c.fluid + c.wall < (1ll << 31) - 1 && c.cells < (1ll << 31) - 1 &&
               c.fluid_entries < (1ll << 40) && c.sims <= (1 << 20), TSPH_ERR_ARG, "capacities too large");
  // slots per row, rounded down to whole chunks of four
  const int64_t kf = (c.fluid_entries / c.fluid) & ~3ll, kw = c.wall > 0 ? (c.wall_entries / c.wall) & ~3ll : 0;
  TSPH_REQUIRE(kf >= 32 && kf <= 256 && (c.wall == 0 || (kw >= 32 && kw <= 256)), TSPH_ERR_ARG,
               "neighbour-list entries must give 32..256 slots per row (fluid %lld, wall %lld)",
               (long long)kf, (long long)kw);
  tsph_batch* B = new tsph_batch();
  B->caps = c;
  const size_t F = (size_t)c.fluid, W = (size_t)c.wall + 1, S = (size_t)c.sims, C = (size_t)c.cells;
  const size_t N = F + W;
  struct { DevBuf* b; size_t bytes; } plan[] = {
      {&B->x[0], 16 * F}, {&B->x[1], 16 * F}, {&B->v[0], 16 * F}, {&B->v[1], 16 * F},
      {&B->rho[0], 8 * F}, {&B->rho[1], 8 * F}, {&B->m, 8 * F}, {&B->fsim, 4 * F},
      {&B->xw, 16 * W}, {&B->vw, 16 * W}, {&B->nw, 16 * W}, {&B->wsim, 4 * W},
      {&B->dps, sizeof(DpSim) * S}, {&B->bs, sizeof(BSim) * S}, {&B->state, S}, {&B->active, S},
      {&B->status, 4 * S}, {&B->vmax, 8 * S}, {&B->bbox, 32 * S}, {&B->total_cells, 8},
      {&B->flags, 8}, {&B->step_counter, 8},
      {&B->cid, 4 * N}, {&B->ccount, 4 * C}, {&B->cfill, 4 * C}, {&B->members, 4 * N},
      {&B->cstart, 8 * (C / kScanTile + 2) * kScanTile}, {&B->cstart_tiles, 8 * (C / kScanTile + 3)},
      {&B->fcnt, 4 * F}, {&B->wcnt, 4 * W},
      {&B->fidx, 4 * ((F + 63) / 64) * 64 * (size_t)kf + 4}, {&B->widx, 4 * ((W + 63) / 64) * 64 * (size_t)kw + 4},
      {&B->wall_p, 8 * W}, {&B->wall_rho, 8 * W}, {&B->reaction, 16 * W}, {&B->acc, 16 * F},
      {&B->lam, 8 * F}, {&B->inv, 32 * F}, {&B->gr, 16 * F}, {&B->normal, 16 * F}, {&B->strength, 8 * F},
      {&B->shift, 16 * F}, {&B->surface, F}, {&B->vol, 8 * F},
      {&B->record, sizeof(BatchRecord)}};
  int rc = TSPH_OK;
  for (auto& p : plan) if (rc == TSPH_OK) rc = p.b->ensure(p.bytes);
  if (rc == TSPH_OK && (hipHostMalloc((void**)&B->host_rec, sizeof(BatchRecord), 0) != hipSuccess ||
                        hipHostMalloc((void**)&B->host_state, S, 0) != hipSuccess ||
                        hipHostMalloc((void**)&B->host_steps, 8 * S, 0) != hipSuccess)) {
    set_error("hipHostMalloc of the step record failed");
    rc = TSPH_ERR_HIP;
  }
  if (rc != TSPH_OK) { tsph_batch_destroy(B); return rc; }
  memset(B->host_rec, 0, sizeof(BatchRecord));
  memset(B->host_state, 0, S);                   // SS_FREE until a step reports otherwise
  memset(B->host_steps, 0, 8 * S);
  // every row starts as a hole (sim -1), every slot free and inactive
  hipError_t e = hipMemsetAsync(B->fsim.ptr, 0xff, 4 * F, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->wsim.ptr, 0xff, 4 * W, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->state.ptr, 0, S, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->active.ptr, 0, S, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->status.ptr, 0, 4 * S, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->vmax.ptr, 0, 8 * S, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->flags.ptr, 0, 8, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->step_counter.ptr, 0, 8, stream);
  if (e == hipSuccess) e = hipMemsetAsync(B->bs.ptr, 0, sizeof(BSim) * S, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) { set_error("batch init: %s", hipGetErrorString(e)); tsph_batch_destroy(B); return TSPH_ERR_HIP; }
  B->fslot.assign(S, Range{0, 0});
  B->wslot.assign(S, Range{0, 0});
  B->used.assign(S, 0);
  B->done_at.assign(S, -1);
  *out = B;
  return TSPH_OK;
}

int tsph_batch_admit(tsph_batch* B, const tsph_dp_case* c, const double* pos, const double* vel,
                     const double* rho, const double* mass, const double* wall, const double* wall_vel,
                     const double* wall_normal, int32_t* slot, hipStream_t stream) {
  TSPH_REQUIRE(B && c && slot, TSPH_ERR_ARG, "NULL argument");
  *slot = -1;
  TSPH_REQUIRE(c->n > 0 && c->n_wall >= 0 && c->steps > 0, TSPH_ERR_ARG, "a case needs fluid rows and steps");
  TSPH_REQUIRE(pos && vel && rho && mass, TSPH_ERR_ARG, "NULL fluid array");
  TSPH_REQUIRE(c->n_wall == 0 || (wall && wall_vel && wall_normal), TSPH_ERR_ARG, "NULL wall array");
  TSPH_REQUIRE(c->dx > 0 && c->sound_speed > 0 && c->rho0 > 0 && c->courant > 0 && c->courant <= 0.3,
               TSPH_ERR_ARG, "dx, c0, rho0 > 0 and 0 < courant <= 0.3 required");
  const double h = 2.0 * c->dx;
  const double dt_max = c->courant * 2.0 * c->dx / c->sound_speed;
  const double rv = 8.0 * c->viscosity + c->alpha * h * c->sound_speed;
  TSPH_REQUIRE(dt_max * rv / 8.0 <= 0.125 * h * h && dt_max * c->diffusion * c->sound_speed <= 0.125 * h,
               TSPH_ERR_ARG, "the adaptive step would exceed the viscous or diffusive limit");
  int s = -1;
  for (size_t k = 0; k < B->used.size(); ++k) if (!B->used[k]) { s = (int)k; break; }
  if (s < 0) { set_error("no free simulation slot"); return TSPH_ERR_OOM; }
  const int64_t fs = find_range(B->fslot, B->used, B->caps.fluid, c->n);
  const int64_t ws = find_range(B->wslot, B->used, B->caps.wall, c->n_wall);
  if (fs < 0 || ws < 0) { set_error("no room for %lld fluid / %lld wall rows", (long long)c->n, (long long)c->n_wall); return TSPH_ERR_OOM; }
  const int side = B->cur;
  const size_t n = (size_t)c->n, nw = (size_t)c->n_wall;
  TSPH_CHECK_HIP(hipMemcpyAsync(B->x[side].as<double>() + 2 * fs, pos, 16 * n, hipMemcpyDeviceToDevice, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(B->v[side].as<double>() + 2 * fs, vel, 16 * n, hipMemcpyDeviceToDevice, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(B->rho[side].as<double>() + fs, rho, 8 * n, hipMemcpyDeviceToDevice, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(B->m.as<double>() + fs, mass, 8 * n, hipMemcpyDeviceToDevice, stream));
  if (nw) {
    TSPH_CHECK_HIP(hipMemcpyAsync(B->xw.as<double>() + 2 * ws, wall, 16 * nw, hipMemcpyDeviceToDevice, stream));
    TSPH_CHECK_HIP(hipMemcpyAsync(B->vw.as<double>() + 2 * ws, wall_vel, 16 * nw, hipMemcpyDeviceToDevice, stream));
    TSPH_CHECK_HIP(hipMemcpyAsync(B->nw.as<double>() + 2 * ws, wall_normal, 16 * nw, hipMemcpyDeviceToDevice, stream));
  }
  k_b_fill_sim<<<nblk((int64_t)n), kBlock, 0, stream>>>(B->fsim.as<int32_t>(), fs, (int64_t)n, s);
  if (nw) k_b_fill_sim<<<nblk((int64_t)nw), kBlock, 0, stream>>>(B->wsim.as<int32_t>(), ws, (int64_t)nw, s);
  k_b_vmax_range<<<1, 1024, 0, stream>>>(B->v[side].as<double>(), fs, (int64_t)n, B->vmax.as<unsigned long long>() + s);
  TSPH_CHECK_LAUNCH();
  unsigned long long vm_enc = 0;
  TSPH_CHECK_HIP(hipMemcpyAsync(&vm_enc, B->vmax.as<unsigned long long>() + s, 8, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  const double vm = ord_dec_host(vm_enc);
  DpSim P = dp_sim_params(c->dx, c->sound_speed, c->reference_speed, c->courant * 2.0 * c->dx / (c->sound_speed + vm),
                          c->rho0, c->gravity, c->viscosity, c->diffusion, c->alpha);
  BSim q{};
  q.cell = 2.0 * h;
  q.courant = c->courant;
  q.budget = c->steps;
  q.steps = 0;
  q.done_step = -1;
  const unsigned char one = SS_ACTIVE, zero = 0;
  const int izero = 0;
  const unsigned long long uzero = 0;
  TSPH_CHECK_HIP(hipMemcpy(B->dps.as<DpSim>() + s, &P, sizeof(DpSim), hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->bs.as<BSim>() + s, &q, sizeof(BSim), hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->status.as<int>() + s, &izero, sizeof(int), hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->vmax.as<unsigned long long>() + s, &uzero, 8, hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->state.as<unsigned char>() + s, &one, 1, hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->active.as<unsigned char>() + s, &one, 1, hipMemcpyHostToDevice));
  (void)zero;
  B->used[s] = 1;
  B->host_state[s] = SS_ACTIVE;                  // until the next step's record says otherwise
  B->fslot[s] = Range{fs, (int64_t)n};
  B->wslot[s] = Range{ws, (int64_t)nw};
  B->done_at[s] = -1;
  B->f_hw = std::max(B->f_hw, fs + (int64_t)n);
  B->w_hw = std::max(B->w_hw, ws + (int64_t)nw);
  *slot = s;
  return TSPH_OK;
}

int tsph_batch_step(tsph_batch* B, int64_t nsteps, hipStream_t stream) {
  TSPH_REQUIRE(B && nsteps >= 0, TSPH_ERR_ARG, "bad arguments");
  if (!B->graph[0] || B->graph_stream != stream) TSPH_TRY(batch_capture(B, stream));
  for (int64_t k = 0; k < nsteps; ++k) {
    TSPH_CHECK_HIP(hipGraphLaunch(B->graph[B->cur], stream));
    B->cur ^= 1;
    B->steps_done++;
  }
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  TSPH_REQUIRE(!(B->host_rec->flags & 1u), TSPH_ERR_OOM, "cell capacity exceeded (grow tsph_batch_caps.cells)");
  return TSPH_OK;
}

int tsph_batch_launch(tsph_batch* B, int32_t sweep_block, int32_t list_block, int32_t xcd_chunks) {
  TSPH_REQUIRE(B, TSPH_ERR_ARG, "batch must not be NULL");
  auto ok = [](int32_t b) { return b == 64 || b == 128 || b == 192 || b == 256; };
  TSPH_REQUIRE(ok(sweep_block) && ok(list_block), TSPH_ERR_ARG,
               "workgroup sizes must be 64, 128, 192 or 256 rows (got %d, %d)", (int)sweep_block,
               (int)list_block);
  TSPH_REQUIRE(xcd_chunks == 0 || xcd_chunks == 1, TSPH_ERR_ARG, "xcd_chunks must be 0 or 1");
  if (B->blk_sweep == sweep_block && B->blk_lists == list_block && B->xcd == xcd_chunks) return TSPH_OK;
  B->blk_sweep = sweep_block;
  B->blk_lists = list_block;
  B->xcd = xcd_chunks;
  for (int side = 0; side < 2; ++side)          // captured again at the next step
    if (B->graph[side]) { (void)hipGraphExecDestroy(B->graph[side]); B->graph[side] = nullptr; }
  return TSPH_OK;
}

int tsph_batch_poll(tsph_batch* B, int32_t* slots, int32_t* states, int32_t max, int32_t* nout) {
  TSPH_REQUIRE(B && nout, TSPH_ERR_ARG, "NULL argument");
  int32_t k = 0;
  for (size_t s = 0; s < B->used.size(); ++s) {
    if (!B->used[s]) continue;
    const unsigned char st = B->host_state[s];
    if (st == SS_DONE || st == SS_FAILED) {
      if (k < max && slots && states) { slots[k] = (int32_t)s; states[k] = st; }
      ++k;
    }
  }
  *nout = k;
  return TSPH_OK;
}

int tsph_batch_read(tsph_batch* B, int32_t s, double* pos, double* vel, double* rho, int64_t* steps,
                    hipStream_t stream) {
  TSPH_REQUIRE(B && s >= 0 && (size_t)s < B->used.size() && B->used[s], TSPH_ERR_ARG, "no such simulation");
  BSim q{};
  TSPH_CHECK_HIP(hipMemcpy(&q, B->bs.as<BSim>() + s, sizeof(BSim), hipMemcpyDeviceToHost));
  // the state of a simulation lives on the side written by its last step:
  // a running one on the current side, a finished one on the side after the
  // step at which it finished (the graph of step k writes side (k + 1) & 1)
  const int side = q.done_step >= 0 ? (int)((q.done_step + 1) & 1) : B->cur;
  const Range f = B->fslot[s];
  if (pos) TSPH_CHECK_HIP(hipMemcpyAsync(pos, B->x[side].as<double>() + 2 * f.start, 16 * f.count, hipMemcpyDeviceToDevice, stream));
  if (vel) TSPH_CHECK_HIP(hipMemcpyAsync(vel, B->v[side].as<double>() + 2 * f.start, 16 * f.count, hipMemcpyDeviceToDevice, stream));
  if (rho) TSPH_CHECK_HIP(hipMemcpyAsync(rho, B->rho[side].as<double>() + f.start, 8 * f.count, hipMemcpyDeviceToDevice, stream));
  if (steps) *steps = q.steps;
  return TSPH_OK;
}

int tsph_batch_release(tsph_batch* B, int32_t s, hipStream_t stream) {
  TSPH_REQUIRE(B && s >= 0 && (size_t)s < B->used.size() && B->used[s], TSPH_ERR_ARG, "no such simulation");
  const unsigned char zero = SS_FREE;
  TSPH_CHECK_HIP(hipMemcpy(B->state.as<unsigned char>() + s, &zero, 1, hipMemcpyHostToDevice));
  TSPH_CHECK_HIP(hipMemcpy(B->active.as<unsigned char>() + s, &zero, 1, hipMemcpyHostToDevice));
  const Range f = B->fslot[s], w = B->wslot[s];
  k_b_fill_sim<<<nblk(std::max<int64_t>(f.count, 1)), kBlock, 0, stream>>>(B->fsim.as<int32_t>(), f.start, f.count, -1);
  if (w.count) k_b_fill_sim<<<nblk(w.count), kBlock, 0, stream>>>(B->wsim.as<int32_t>(), w.start, w.count, -1);
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  B->host_state[s] = SS_FREE;
  B->used[s] = 0;
  B->fslot[s] = Range{0, 0};
  B->wslot[s] = Range{0, 0};
  return TSPH_OK;
}

int tsph_batch_stats(tsph_batch* B, int64_t* counters) {
  TSPH_REQUIRE(B && counters, TSPH_ERR_ARG, "NULL argument");
  int64_t live = 0, fl = 0, wl = 0;
  for (size_t s = 0; s < B->used.size(); ++s)
    if (B->used[s]) { ++live; fl += B->fslot[s].count; wl += B->wslot[s].count; }
  counters[0] = B->steps_done;
  counters[1] = live;
  counters[2] = fl;
  counters[3] = wl;
  return TSPH_OK;
}

}  // extern "C"
