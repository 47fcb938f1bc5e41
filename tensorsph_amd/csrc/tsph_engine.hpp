// Resident stepper state shared by the list-based (tsph_engine.hip) and the
// cell-tile (tsph_tiles.hip) sweeps.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

#include "tsph_bin.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"

namespace tsph {

// device flag block
struct SimFlags {
  unsigned long long maxes[MM_NMAX];
  unsigned long long escaped;
  unsigned long long stale[2];      // slot k&1 = input of step k, slot (k+1)&1 = output
  unsigned long long mins[MN_NMIN]; // ordered keys, atomicMin
  unsigned long long lds_overflow;  // tiles that ran from HBM (block larger than the LDS budget)
  // time step of the coming step, decided on the device after every step
  double dt;                        // the step size variable of next_dt
  double step_dt;                   // what the coming step integrates with (0: final sample)
  double elapsed;                   // simulated time at the start of the coming step
  double last_dt;                   // the previous step's step_dt (op.dt before the first)
  unsigned long long final;         // adaptive: the coming step is the final sample
};

// adaptive step control (droplet_impact.py:495-552, 1674-1763)
struct DtCtl {
  int adaptive;
  int free_last;                    // 1: the last step is not shortened onto target_time
  double dt_cfg, target_time, tol;
  double courant_num;               // max_courant * h_factor * dx, as the reference forms it
  double c0, dt_limit, dt_safety, dt_growth;
  double t0, dt0;                   // time state of a resumed run: elapsed, step-size variable
};

// one entry of the host ring, packed on the device after every step and
// copied to pinned memory in one transfer
struct StepRecord {
  unsigned long long stale_in, stale_out, escaped;
  unsigned long long maxes[MM_NMAX];
  unsigned long long mins[MN_NMIN];
  double sums[MS_NSUM];
  double loads[6];                       // obstacle force and moment of this step
  double clamp_imp[3];                   // cumulative obstacle-clamp impulse
  unsigned long long clamp_count;        // cumulative clamped particles
  double elapsed, last_dt, step_dt, dt;  // time state at the start of this step
  unsigned long long final;              // adaptive: this step was the final sample
};

constexpr int kRing = 4;
// doubles per meter row: step, sums, maxima, minima, time state
constexpr int kMeterWidth = 1 + MS_NSUM + MM_NMAX + MN_NMIN + 4;

// one workgroup's tile: its cell-tile id and its fluid (or wall) rows
struct TileRec {
  int64_t tile;
  int32_t start, end;
};

}  // namespace tsph

struct tsph_sim {
  tsph_sim_params prm{};
  int dim = 2, prec = 0;
  int64_t n = 0, nw = 0, nt = 0;
  tsph_ctx* ctx = nullptr;
  tsph::DevBuf pos[2], vel[2], mass, mass2, orig, orig2;
  tsph::DevBuf rho, p, pr2, vol, rho_raw, rho_lo, grad, acc, anchor;
  tsph::DevBuf aux;                 // FP32: Vec4 (m, q_hi, m / rho, q_lo) per row, q = p / rho^2
  tsph::DevBuf p_lo, gsum;          // FP32: p's dropped part per row; sum_j m_j grad W_ij per fluid row
  tsph::DevBuf closure;             // free-surface pressure closure: two pressure and two flag buffers
  tsph::DevBuf shift_buf, shift_tmp;  // particle shifting: C, sorted C, shift vectors, counters; sort scratch
  double last_speed = 0.0;          // largest speed at the start of the last step run
  double last_step_dt = 0.0;        // the size of the last step run
  bool rebuild_forced = false;      // a correction re-binned uncounted: the next step re-bins
  int64_t escaped_extra = 0;        // boundary crossings of the position corrections
  tsph::DevBuf flags, block_sums, block_max, sums_out, rec_dev;
  tsph::DevBuf new2old, cells_new;
  // binning of the slot-list layouts: the fixed grid (set at a full build,
  // which also bins the walls) and the counting sort's buffers
  tsph::BinGrid bgrid{};
  bool bgrid_ok = false;
  tsph::DevBuf bin_cnt, bin_start, bin_part, bin_keys, bin_tmp, bin_cells, bin_flag, bin_bounds;
  unsigned int* bin_outside = nullptr;   // pinned: a row left the grid at the last binning
  tsph::BuildStats* build_stats = nullptr;  // pinned: the last halo-list build's report
  tsph::DevBuf cell_R;                    // int4 per cell: fluid / wall row ranges (cell walk)
  int64_t ncells_dense = 0, cell_ext[3] = {1, 1, 1}, cell_stride[3] = {0, 0, 0};
  tsph::DevBuf slot_idx, slot_cnt, slot_max;   // per-row lists in fixed slots (NbSlots)
  tsph::DevBuf slot_posf;                      // FP32 build positions (list prefilter)
  int64_t slot_K = 0;
  int64_t slot_K_next = -1;                    // >= 0: the width the next build starts from
  tsph::StepRecord* ring = nullptr;       // pinned
  hipEvent_t ev[tsph::kRing] = {};        // the step's record is in the ring
  hipEvent_t pev[tsph::kRing] = {};       // the step's record is packed (stepping stream)
  hipStream_t rec_stream = nullptr;       // copies the records to the ring
  // per-sweep timing (HIP events around each sweep on the stepping stream)
  int timing = 0;                   // > 0: time the sweeps of every timing-th step
  bool tstep = false;               // the step being queued is timed
  bool timed[tsph::kRing] = {};
  hipEvent_t tev[tsph::kRing][4] = {};
  double sweep_ms[3] = {0, 0, 0};
  int64_t sweep_count = 0;
  int64_t pairs_2h = -1;
  int64_t ring_step[tsph::kRing];
  int cur = 0;
  bool rebuild_pending = false;  // the last step of a run went stale: rebuild before the next step
  int64_t steps_done = 0, rebuilds = 0, escaped_total = 0;
  bool built = false;
  bool finished = false;                  // an adaptive run has done its final sample
  // halo mode (2-D slot lists, tiles == 2): per tile of 256 fluid rows, the
  // rows its lists name (staged in LDS by the density and force sweeps) and
  // the lists as 16-bit halo positions
  tsph::DevBuf halo_rows, halo_cnt, halo_idx, halo_hmax;
  int64_t halo_tiles = 0;        // tiles of the build
  int halo_max = 0;              // halo rows the sweeps stage (LDS rows to reserve)
  bool halo_ok = false;          // the halo of the current build is in place
  // cell-tile mode (2-D): rows sorted tile-major, dense per-cell row ranges
  int tiles = 0;                 // 1: LDS cell-tile force sweep, 0: slot-list sweeps, 2: slot lists + halos
  int tile_dim[3] = {4, 4, 1};
  int64_t ext[3] = {1, 1, 1};
  int64_t ntile[3] = {1, 1, 1};
  int64_t ntiles_f = 0, ntiles_w = 0;
  tsph::DevBuf ranges;           // int32 4 x ncells: fluid start/end, wall start/end
  tsph::DevBuf tiles_f, tiles_w; // TileRec
  tsph::DevBuf tkeys, tkeys_sorted, tidx, tflags;
  int64_t lds_overflow = 0;
  // host-side meter log
  std::vector<double> rows;         // per recorded step: step, sums..., maxes..., mins...
  // obstacle loads (dim 3): per-block partials, reduced loads, clamp accumulators
  int obstacle = 0;
  tsph::DevBuf load_part, loads_out, clamp_acc;
  std::vector<double> load_rows;    // per completed step: step, F, M, clamp impulse, clamp count
  double peak_rho = 0.0, peak_speed = 0.0;
  size_t elt = 8;
};


namespace tsph {
int sim_consume(tsph_sim* sim, int64_t k, bool meters, bool* stale_after, bool* skipped = nullptr);
__global__ void k_reset_meters(SimFlags* fl, int slot_in);
__global__ void k_reduce_sums(const double* part, int64_t nblocks, double* out, SimFlags* fl, int slot_in);
template <int D, typename S> int sim_rebuild_tiles(tsph_sim* sim, hipStream_t stream);
int tile_count_pairs(tsph_sim* sim, double radius, unsigned long long* cnt, hipStream_t stream);
template <int D, typename S> int sim_queue_step_tiles(tsph_sim* sim, int64_t k, bool meters, hipStream_t stream);
template <int D>
int build_sim_slots(tsph_sim* sim, const double* npos, double rc, const int64_t* ext, const int64_t* stride,
                    hipStream_t stream);
template <int D, typename S> int queue_slot_density_wall(tsph_sim* sim, int64_t k, hipStream_t stream);
}  // namespace tsph
