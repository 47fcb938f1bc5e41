// Resident WCSPH stepper: the per-step loop of droplet_impact.run
// (droplet_impact.py:1674-1763), dam_break.run_case (dam_break.py:374-465)
// and, in 3-D, kleefsman_obstacle_3d.run_case (:644-706), with the particle
// state kept in HBM between steps.
//
// Layout.  Fluid rows [0, n) then wall rows [n, n + nw), positions and
// velocities FP64 (nt, D), ping-ponged between two buffers.  At every
// rebuild the particles are binned into cells of edge 2h + skin and the rows
// re-ordered by the Morton key of their cell (fluid among fluid, wall among
// wall), so a particle's neighbours sit close to it in memory; orig[] maps
// every row back to the caller's index.  A dense table of per-cell row
// ranges replaces per-particle lists: each sweep walks the 3^D cells around
// its target in the order the reference-format lists would list them
// (NbCells, tsph_cells.hpp).
//
// Verlet skin without a per-step host round trip.  The binning at 2h + skin
// stays valid while no particle has moved more than skin/2 since it was made
// (neighbor_search.py:219-227).  Step k's force sweep
// raises a device flag when its new positions break that bound; every sweep
// of step k+1 reads the flag first and, when set, does nothing (the force
// sweep copies the state through).  The host learns of the flag one step
// later from a pinned-memory copy, rebuilds, and re-runs the skipped steps,
// so the schedule of rebuilds is exactly the reference's while the host
// never blocks the GPU queue.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tsph_cells.hpp"
#include "tsph_common.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"
#include "tsph_engine.hpp"
#include "tsph_halo.hpp"

#ifndef TSPH_HALO_CH
#define TSPH_HALO_CH 1
#endif

namespace tsph {

TSPH_D double ord_dec(unsigned long long u) {
  const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  return __longlong_as_double((long long)b);
}

// Packs the record of the step just queued and, unless it was skipped,
// advances the time state: elapsed += step_dt, and with the adaptive step
// dt = next_dt(dt, max |v| at the step's start) (droplet_impact.py:536-552),
// the next step_dt = min(dt, target - elapsed), and the step that starts at the
// target (within tol) is the final sample, integrated with step_dt = 0.
// rec: the step's record slot (device).  The force sweep's per-block
// records (blk, nblk of them; NULL when it used atomics) are reduced here first: the displacement, speed and density
// maxima and the boundary hits of the step.  A skipped (stale) step wrote
// none.  kPackThreads threads.
constexpr int kPackThreads = 1024;
__global__ void __launch_bounds__(kPackThreads) k_pack_record(
    SimFlags* fl, const double* __restrict__ sums, int slot_in, int with_sums, StepRecord* __restrict__ rec,
    const double* __restrict__ loads, const double* __restrict__ clamp,
    const unsigned long long* __restrict__ blk, int64_t nblk, DtCtl ctl) {
  const int t = threadIdx.x;
  const bool fresh = blk && !fl->stale[slot_in];
  unsigned long long key[3] = {0ull, 0ull, 0ull}, hits = 0ull;
  if (fresh) {
    // four records per thread and round, their loads issued together
    for (int64_t b0 = t; b0 < nblk; b0 += 4 * kPackThreads) {
      ulonglong2 a[4], c[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t b = b0 + q * kPackThreads;
        const int64_t bb = b < nblk ? b : b0;
        a[q] = reinterpret_cast<const ulonglong2*>(blk)[2 * bb];
        c[q] = reinterpret_cast<const ulonglong2*>(blk)[2 * bb + 1];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        key[0] = a[q].x > key[0] ? a[q].x : key[0];
        key[1] = a[q].y > key[1] ? a[q].y : key[1];
        key[2] = c[q].x > key[2] ? c[q].x : key[2];
        hits += b0 + q * kPackThreads < nblk ? c[q].y : 0ull;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const unsigned long long y = __shfl_xor(key[q], o, 64);
      key[q] = y > key[q] ? y : key[q];
    }
    hits += __shfl_xor(hits, o, 64);
  }
  __shared__ unsigned long long part[4][kPackThreads / 64];
  if ((t & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) part[q][t >> 6] = key[q];
    part[3][t >> 6] = hits;
  }
  __syncthreads();
  if (t == 0 && fresh) {
    for (int w = 1; w < kPackThreads / 64; ++w) {
#pragma unroll
      for (int q = 0; q < 3; ++q) part[q][0] = part[q][w] > part[q][0] ? part[q][w] : part[q][0];
      part[3][0] += part[3][w];
    }
    if (part[0][0] > fl->maxes[MM_DISP]) fl->maxes[MM_DISP] = part[0][0];
    fl->escaped += part[3][0];
  }
  __syncthreads();
  if (t == 0) { rec->stale_in = fl->stale[slot_in]; rec->stale_out = fl->stale[slot_in ^ 1]; rec->escaped = fl->escaped; }
  if (t < MM_NMAX && t != MM_PEAK_SPEED && t != MM_PEAK_RHO) rec->maxes[t] = fl->maxes[t];
  if (t < MN_NMIN) rec->mins[t] = fl->mins[t];
  if (t < MS_NSUM) rec->sums[t] = with_sums ? sums[t] : 0.0;
  if (t < 6) rec->loads[t] = loads ? loads[t] : 0.0;
  if (t < 3) rec->clamp_imp[t] = clamp ? clamp[t] : 0.0;
  if (t == 0) {
    rec->clamp_count = clamp ? ((const unsigned long long*)(clamp + 3))[0] : 0ull;
    rec->elapsed = fl->elapsed; rec->last_dt = fl->last_dt; rec->step_dt = fl->step_dt; rec->dt = fl->dt;
    rec->final = fl->final;
    // the step's largest speed and density, read and cleared in one atomic
    // each (a plain scalar load followed by an independent store of 0 is not
    // ordered against it); the host keeps the run's peaks
    // (the record path brings them in its per-block records, the cell-tile
    // path leaves them in the flag block)
    unsigned long long kv = 0ull, kr = 0ull;
    if (blk) {
      kv = fresh ? part[1][0] : 0ull;
      kr = fresh ? part[2][0] : 0ull;
    } else {
      kv = atomicExch(&fl->maxes[MM_PEAK_SPEED], 0ull);
      kr = atomicExch(&fl->maxes[MM_PEAK_RHO], 0ull);
    }
    rec->maxes[MM_PEAK_SPEED] = kv;
    rec->maxes[MM_PEAK_RHO] = kr;
    const double vmax = ord_dec(kv);
    if (!fl->stale[slot_in] && !fl->final) {
      fl->elapsed += fl->step_dt;
      fl->last_dt = fl->step_dt;
      if (ctl.adaptive) {
        const double v = vmax > 0.0 ? vmax : 0.0;
        const double acoustic = ctl.courant_num / (ctl.c0 + v);
        const double target = ctl.dt_safety * (acoustic < ctl.dt_limit ? acoustic : ctl.dt_limit);
        const double grown = ctl.dt_growth >= 1.0 ? fl->dt * ctl.dt_growth : target;
        double dt = target < grown ? target : grown;
        dt = dt < ctl.dt_cfg ? dt : ctl.dt_cfg;
        fl->dt = dt;
        const double left = ctl.target_time - fl->elapsed;
        fl->final = fl->elapsed >= ctl.target_time - ctl.tol ? 1ull : 0ull;
        fl->step_dt = fl->final ? 0.0 : (ctl.free_last || dt < left ? dt : left);
      }
    }
  }
}

// a run keeps per-step load rows when it has obstacle stations or the clamp
static inline bool load_rows(const tsph_sim* sim) { return sim->obstacle || sim->prm.obstacle_clamp; }

// the per-row slot lists built at the last rebuild
static inline NbSlots sim_slots(const tsph_sim* sim) {
  NbSlots l{};
  l.idx = sim->slot_idx.as<int32_t>();
  l.cnt = sim->slot_cnt.as<int32_t>();
  l.K = sim->slot_K;
  l.nf = sim->n;
  return l;
}

// the cell-walk neighbour source over the rows binned at the last rebuild
static inline NbCells sim_cells(const tsph_sim* sim) {
  NbCells c{};
  c.R = sim->cell_R.as<int4>();
  c.cells = sim->cells_new.as<int64_t>();
  for (int d = 0; d < 3; ++d) { c.ext[d] = sim->cell_ext[d]; c.stride[d] = sim->cell_stride[d]; }
  c.periodic = 0;
  return c;
}

// Obstacle loads of one step: the rhs_3d reaction of every obstacle station
// (the k_wall_loads formula, dam_break_3d.py:410-422) and its moment about the
// centre, reduced per block (partials in a fixed order, no float atomics).
template <int D, typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_loads(WcsphArgs<S> a, NB nb,
                                                       const int64_t* __restrict__ orig, int64_t obs_lo,
                                                       int64_t obs_hi, double c0, double c1, double c2,
                                                       SimFlags* fl, int slot_in, double* __restrict__ part) {
  if (fl->stale[slot_in]) return;                // uniform over the grid
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double out[6] = {0, 0, 0, 0, 0, 0};
  if (w < a.nw) {
    const int64_t iw = a.n + w;
    const int64_t o = orig[iw];
    if (o >= obs_lo && o < obs_hi) {
      double xw[D], vw[D];
#pragma unroll
      for (int d = 0; d < D; ++d) { xw[d] = a.pos[iw * D + d]; vw[d] = a.vel[iw * D + d]; }
      const S prw = a.pr2[iw], mw = a.mass[iw], volw = a.vol[iw];
      Sum<S> r[D];
      nb.template visit<D>(iw, false, [&](int64_t f) {
        S df[D], r2 = S(0);
#pragma unroll
        for (int d = 0; d < D; ++d) { df[d] = (S)(a.pos[f * D + d] - xw[d]); r2 += df[d] * df[d]; }
        if (r2 > a.skip2) return;
        S wgt, g;
        cubic_wg<S>(sqrt_s<S>(r2), a.k, wgt, g);
        const S mf = a.mass[f];
        const S term = a.pr2[f] + prw;
        S dg = S(0);
#pragma unroll
        for (int d = 0; d < D; ++d) dg += df[d] * (df[d] * g);
        const S relu = -dg > S(0) ? -dg : S(0);
        const S lw = relu / (r2 + a.eta2);
        const S cv = -(a.nu2 * a.friction) * (volw * lw);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const S pp = (mf * mw) * (term * (df[d] * g));
          const S visc_f = cv * (S)(a.vel[f * D + d] - vw[d]);
          r[d].add(pp - mf * visc_f);
        }
      });
      double F[3] = {0, 0, 0}, R[3] = {0, 0, 0};
      const double cc[3] = {c0, c1, c2};
#pragma unroll
      for (int d = 0; d < D; ++d) { F[d] = (double)r[d].get(); R[d] = xw[d] - cc[d]; }
      out[0] = F[0]; out[1] = F[1]; out[2] = F[2];
      out[3] = R[1] * F[2] - R[2] * F[1];
      out[4] = R[2] * F[0] - R[0] * F[2];
      out[5] = R[0] * F[1] - R[1] * F[0];
    }
  }
  __shared__ double red[6][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = out[k];
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) v += red[threadIdx.x][q];
    part[(int64_t)blockIdx.x * 6 + threadIdx.x] = v;
  }
}

// Sums the per-block partials of k_step_loads in a fixed order (thread t takes
// blocks t, t + 256, ..., then a fixed tree over the threads): one block of 256.
__global__ void __launch_bounds__(256) k_reduce_loads(const double* __restrict__ part, int64_t nblocks,
                                                      double* __restrict__ out, const SimFlags* fl, int slot_in) {
  if (fl->stale[slot_in]) return;
  __shared__ double red[6][256];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += part[b * 6 + k];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}



// ------------------------------------------------------------------ step sweeps
// Row i's fluid state from its density, known to FP64 accuracy.  FP64 runs:
// the power-form Tait EOS and p / rho^2 of the stored values.  FP32 runs
// evaluate the EOS and q = p / rho^2 in FP64 too and keep q as q_hi + q_lo
// (about 2^-48 relative) and p as p + p_lo.  The force sweep forms the
// pressure contrast q_j - q_i from both parts: at depth H a row's q is about
// H / h times the contrast it carries to a neighbour, so a q rounded to FP32
// would put that factor times 2^-24 into the force.  The wall sweep
// extrapolates p to FP64 accuracy from p + p_lo.
template <typename S>
TSPH_D void store_fluid_state(const WcsphArgs<S>& a, int64_t i, S rho, double rho_d) {
  if constexpr (sizeof(S) == 4) {
    const double p_d = tait<double>(rho_d, a.rho0d, a.bulkd, a.gammad, a.ratio_floord);
    const double q_d = p_d / (rho_d * rho_d);
    const S p = (S)p_d, q = (S)q_d, ql = (S)(q_d - (double)q);
    const S vol = a.mass[i] / rho;
    a.p[i] = p;
    if (a.p_lo) a.p_lo[i] = (S)(p_d - (double)p);
    a.pr2[i] = q;
    a.vol[i] = vol;
    if (a.aux) a.aux[i] = Vec4<S>{a.mass[i], q, vol, ql};
  } else {
    const S p = tait_hi<S>(rho_d, a.rho0d, a.bulk, a.gamma, a.ratio_floor);
    const S pr2 = p / (rho * rho), vol = a.mass[i] / rho;
    a.p[i] = p;
    a.pr2[i] = pr2;
    a.vol[i] = vol;
  }
}

// r = sqrt(r2) and 1 / r to about 5e-14 relative from the FP32 hardware
// reciprocal square root and one Newton step in FP64.  r2 below 1e-30 is
// taken as 1e-30 for 1 / r (finite, 1e15), so r = r2 / sqrt(1e-30) <= 1e-15
// there and r = 0 at r2 = 0, without selects.
TSPH_D void sqrt_rinv_newton(double r2, double& r, double& rinv) {
  const double rc = fmax(r2, 1.0e-30);
  const double y0 = (double)__builtin_amdgcn_rsqf((float)rc);
  rinv = y0 * (1.5 - (0.5 * rc) * (y0 * y0));
  r = r2 * rinv;
}

// HALO: the tile's halo rows (positions, masses) are staged in LDS and the
// lists of halo positions are evaluated from there (2-D slot lists); a tile
// without a halo reads the global arrays.  Same terms in the same order
// either way.
template <int D, typename S, typename NB, bool HALO = false, bool SHEP = false>
__global__ void __launch_bounds__(kBlock) k_step_density(WcsphArgs<S> a, NB nbr, SimFlags* fl,
                                                         int slot_in, int write_eos, HaloView hv = {},
                                                         NbHalo hl = {}) {
  // FP32 runs: masses as hi + lo (a.mass_lo) and the pressure split's G_i
  // (a.gsum) are always present
  constexpr bool kSplit = sizeof(S) == 4;
  if (fl->stale[slot_in]) return;
  const int64_t nb = gridDim.x;
  const int64_t tile = xcd_block(blockIdx.x, nb);
  const int64_t i = tile * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) fl->maxes[MM_DISP] = 0;
  int H = -1;
  extern __shared__ __align__(16) unsigned char halo_lds[];
  double* hx = reinterpret_cast<double*>(halo_lds);
  double* hm = nullptr;
  if constexpr (HALO) {
    H = tile < hv.ntiles ? hv.cnt[tile].x : -1;
    if (H > hv.max_rows) H = -1;                     // more rows than this launch staged: global path
    if (H >= 0) {
      hm = hx + (size_t)H * D;
      // every load of the staging is issued before the first store (up to
      // halo_cap / kBlock rows per thread)
      const uint32_t* hr = hv.rows + tile * halo_cap<D>();
      constexpr int QS = halo_cap<D>() / kBlock;
      double px[QS][D], pm[QS];
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s = threadIdx.x + q * kBlock;
        const uint32_t j = hr[s < H ? s : 0];
#pragma unroll
        for (int d = 0; d < D; ++d) px[q][d] = a.pos[(int64_t)j * D + d];
        pm[q] = (double)a.mass[j];
        if constexpr (kSplit) pm[q] += (double)a.mass_lo[j];
      }
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s = threadIdx.x + q * kBlock;
        if (s < H) {
#pragma unroll
          for (int d = 0; d < D; ++d) hx[s * D + d] = px[q][d];
          hm[s] = pm[q];
        }
      }
      __syncthreads();
    }
  }
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> acc, kt;
  // FP32 runs also sum G_i = sum_j m_j grad W_ij over the same entries, in
  // FP64: the force sweep adds the pressure terms as -2 q_i G_i plus the
  // contrast terms -m_j (q_j - q_i) grad W_ij (the same sum, regrouped)
  double gs[D] = {};
  // one entry: its position and mass, fetched from the halo or the arrays
  auto entry = [&](const double (&pj)[D], double mj, bool keep) {
    double r2 = 0.0, df[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      df[d] = xi[d] - pj[d];
      r2 += df[d] * df[d];
    }
    keep = keep && !(r2 > a.skip2d);
    // FP32-compensated runs take r from the refined reciprocal square root
    // (5e-14 relative) and W from the branch-free form; FP64 runs keep the
    // correctly rounded square root and cubic_wg
    double w, g;
    if constexpr (kSplit) {
      double r, rinv, dw;
      sqrt_rinv_newton(r2, r, rinv);
      cubic_w_dw_select(r, a.kd, w, dw);
      // an entry outside the support adds +0 to every sum, as add_if leaves it
      const double mk = keep ? mj : 0.0;
      const double gm = mk * (dw * rinv);
#pragma unroll
      for (int d = 0; d < D; ++d) gs[d] = fma(gm, df[d], gs[d]);
      acc.add(mk * w);
    } else {
      cubic_wg_select<double>(sqrt(r2), a.kd, w, g);
      acc.add_if(keep, mj * w);
    }
    if constexpr (SHEP) kt.add_if(keep, w);
  };
  if (HALO && H >= 0) {
    hl.template visit_chunks<D, TSPH_HALO_CH>(i, [&](const int32_t* l4, const bool* keep4) {
      // the chunk's LDS reads first, then its arithmetic
      double pj4[4][D], mj4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int d = 0; d < D; ++d) pj4[u][d] = hx[l4[u] * D + d];
        mj4[u] = hm[l4[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) entry(pj4[u], mj4[u], keep4[u]);
    });
  } else {
    nbr.template visit_masked<D>(i, true, [&](int64_t j, bool keep) {
      const uint32_t jb = (uint32_t)j;
      const double* p = at_bytes(a.pos, jb * (uint32_t)(D * sizeof(double)));
      double pj[D];
#pragma unroll
      for (int d = 0; d < D; ++d) pj[d] = p[d];
      double mj = (double)*at_bytes(a.mass, jb * (uint32_t)sizeof(S));
      if constexpr (kSplit) mj += (double)*at_bytes(a.mass_lo, jb * (uint32_t)sizeof(S));
      entry(pj, mj, keep);
    });
  }
  double mi = (double)a.mass[i];
  if constexpr (kSplit) mi += (double)a.mass_lo[i];
  double rho_d = hi(acc) + mi * a.kd.w0d;
  rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  if constexpr (SHEP) {                 // Shepard renormalisation, droplet_impact.py:954-964
    double t = hi(kt) + a.kd.w0d;
    t = t > 1.0e-12 ? t : 1.0e-12;
    rho_d = rho_d * (a.lattice_sum / t);
    rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  }
  const S rho = store_rho<S>(rho_d, a.rho, a.rho_lo, i);
  if constexpr (kSplit) {
#pragma unroll
    for (int d = 0; d < D; ++d) a.gsum[i * D + d] = (S)gs[d];
  }
  if (write_eos) store_fluid_state<S>(a, i, rho, rho_d);
}

// the density sweep with or without the Shepard factor (a compile-time
// switch: the sum of W it needs costs the default sweep nothing)
template <int D, typename S, bool HALO>
void launch_density(int shepard, unsigned grid, size_t lds, hipStream_t stream, const WcsphArgs<S>& a,
                    const NbSlots& nbr, SimFlags* fl, int slot, int write_eos, const HaloView& hv,
                    const NbHalo& hl) {
  if (shepard)
    k_step_density<D, S, NbSlots, HALO, true><<<grid, kBlock, lds, stream>>>(a, nbr, fl, slot, write_eos, hv, hl);
  else
    k_step_density<D, S, NbSlots, HALO, false><<<grid, kBlock, lds, stream>>>(a, nbr, fl, slot, write_eos, hv, hl);
}

template <int D, typename S, int FORM, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_diffusion(WcsphArgs<S> a, NB nbr,
                                                           const S* __restrict__ rho_in,
                                                           const S* __restrict__ grad, SimFlags* fl,
                                                           int slot_in) {
  if (fl->stale[slot_in]) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  const S rhoi = rho_in[i];
  S gi[D];
  if (FORM == 1) {
#pragma unroll
    for (int d = 0; d < D; ++d) gi[d] = grad[i * D + d];
  }
  Sum<S> rate;
  nbr.template visit<D>(i, false, [&](int64_t j) {
    S df[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)(xi[d] - a.pos[j * D + d]); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) return;
    S w, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
    S radial = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) radial += df[d] * (df[d] * g);
    radial = -radial;
    S contrast = rhoi - rho_in[j];
    if (FORM == 1) {
      S proj = S(0);
#pragma unroll
      for (int d = 0; d < D; ++d) proj += (gi[d] + grad[j * D + d]) * df[d];
      contrast = contrast - S(0.5) * proj;
    }
    rate.add(-a.diff_scale * (a.mass[j] / rho_in[j]) * (contrast * radial / (r2 + S(1e-12))));
  });
  double rho_d = (double)rhoi + (a.rho_lo ? (double)a.rho_lo[i] : 0.0) + (double)(S)fl->step_dt * hi(rate);
  rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  const S rho = (S)rho_d;
  a.rho[i] = rho;
  store_fluid_state<S>(a, i, rho, rho_d);
}

template <int D, typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_gradient(WcsphArgs<S> a, NB nbr,
                                                          S* __restrict__ grad_out, SimFlags* fl,
                                                          int slot_in) {
  if (fl->stale[slot_in]) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  const S rhoi = a.rho[i];
  Sum<S> G[D], M[D * D];
  nbr.template visit<D>(i, false, [&](int64_t j) {
    S df[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)(xi[d] - a.pos[j * D + d]); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) return;
    S w, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
    const S vj = a.mass[j] / a.rho[j];
    const S diff = a.rho[j] - rhoi;
#pragma unroll
    for (int r = 0; r < D; ++r) {
      G[r].add((vj * diff) * (df[r] * g));
#pragma unroll
      for (int c = 0; c < D; ++c) M[r * D + c].add(vj * (-(df[r] * (df[c] * g))));
    }
  });
  double m[9], gv[3], out[3];
  double tr = 0.0;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    gv[r] = (double)G[r].get();
#pragma unroll
    for (int c = 0; c < D; ++c) m[r * D + c] = (double)M[r * D + c].get();
    tr += m[r * D + r];
  }
  if (tr <= 0.0) {
#pragma unroll
    for (int d = 0; d < D; ++d) grad_out[i * D + d] = S(0);
    return;
  }
  regularised_solve<D>(m, gv, (double)a.eig_floor, out);
#pragma unroll
  for (int d = 0; d < D; ++d) grad_out[i * D + d] = (S)out[d];
}

template <int D, typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_wall(WcsphArgs<S> a, NB nbr, SimFlags* fl,
                                                      int slot_in) {
  if (fl->stale[slot_in]) return;
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nw) return;
  const int64_t iw = a.n + w;
  double xw[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xw[d] = a.pos[iw * D + d];
  if constexpr (sizeof(S) == 4) {
    // FP32 runs: the extrapolation and the wall EOS in FP64 from p + p_lo, so
    // that the wall rows' q = q_hi + q_lo carries the pressure contrast to
    // their fluid neighbours as the fluid rows' q does (store_fluid_state)
    Sum<double> num, den;
    const double gy = -(double)a.gravity;
    // four entries at a time: their loads first, then the terms in list order
    nbr.template visit_chunks<D, 1>(iw, false, [&](const int32_t* j4, const bool* keep4) {
      double pf[4][D], ps[4], rf[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t f = j4[u];
#pragma unroll
        for (int d = 0; d < D; ++d) pf[u][d] = a.pos[f * D + d];
        ps[u] = (double)a.p[f] + (double)a.p_lo[f];
        rf[u] = (double)a.rho[f];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        double rw[D], r2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) { rw[d] = xw[d] - pf[u][d]; r2 += rw[d] * rw[d]; }
        const bool keep = keep4[u] && !(r2 > a.skip2d);
        double r, rinv;
        sqrt_rinv_newton(r2, r, rinv);
        const double wgt = cubic_w_select(r, a.kd);
        const double src = gy != 0.0 ? ps[u] + rf[u] * (gy * rw[1]) : ps[u];
        num.add_if(keep, src * wgt);
        den.add_if(keep, wgt);
      }
    });
    const double dd = den.get() > 1e-30 ? den.get() : 1e-30;
    double pw = num.get() / dd;
    pw = pw > 0.0 ? pw : 0.0;
    double rwv = a.rho0d * pow(pw / a.bulkd + 1.0, a.inv_gammad);
    rwv = rwv > a.rho0d ? rwv : a.rho0d;
    const double q_d = pw / (rwv * rwv);
    const S rs = (S)rwv, ps = (S)pw, q = (S)q_d, vol = a.mass[iw] / rs;
    a.p[iw] = ps;
    if (a.p_lo) a.p_lo[iw] = (S)(pw - (double)ps);
    a.rho[iw] = rs;
    a.pr2[iw] = q;
    a.vol[iw] = vol;
    if (a.aux) a.aux[iw] = Vec4<S>{a.mass[iw], q, vol, (S)(q_d - (double)q)};
    return;
  }
  Sum<S> num, den;
  nbr.template visit<D>(iw, false, [&](int64_t f) {
    S rw[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { rw[d] = (S)(xw[d] - a.pos[f * D + d]); r2 += rw[d] * rw[d]; }
    if (r2 > a.skip2) return;
    S wgt, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, wgt, g);
    S src = a.p[f];
    if (a.gravity != S(0)) src = src + a.rho[f] * (-a.gravity * rw[1]);
    num.add(src * wgt);
    den.add(wgt);
  });
  const S dd = den.get() > S(1e-30) ? den.get() : S(1e-30);
  S pw = num.get() / dd;
  pw = pw > S(0) ? pw : S(0);
  S rwv = a.rho0 * pow_s<S>(pw / a.bulk + S(1), a.inv_gamma);
  rwv = rwv > a.rho0 ? rwv : a.rho0;
  const S pr2 = pw / (rwv * rwv), vol = a.mass[iw] / rwv;
  a.p[iw] = pw;
  a.rho[iw] = rwv;
  a.pr2[iw] = pr2;
  a.vol[iw] = vol;
  if (a.aux) a.aux[iw] = Vec4<S>{a.mass[iw], pr2, vol, S(0)};
}

// Boundary-consistent density (droplet_impact.py:1090-1112): the wall rows'
// mass scaled by rho_w / rho0 enters the fluid summation density, then the
// fluid EOS; the wall sweep runs again after it.
template <int D, typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_wall_mass(WcsphArgs<S> a, NB nbr, SimFlags* fl, int slot_in) {
  if (fl->stale[slot_in]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> corr;
  nbr.template visit<D>(i, true, [&](int64_t j) {
    if (j < a.n) return;
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double df = xi[d] - a.pos[j * D + d];
      r2 += df * df;
    }
    if (r2 > a.skip2d) return;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    const double mw = (double)a.mass[j];
    corr.add((mw * ((double)a.rho[j] / a.rho0d) - mw) * w);
  });
  double rho_d = (double)a.rho[i] + hi(corr);
  rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  const S rho = (S)rho_d;
  a.rho[i] = rho;
  store_fluid_state<S>(a, i, rho, rho_d);
}

// Free-surface pressure closure (droplet_impact.py:429-492) on the stepper:
// fluid rows below threshold * rho0 take the kernel-weighted extrapolation of
// their fluid neighbours' pressure, from donors known at the pass start
// ("interior", several passes) or all neighbours ("all", one pass).  The
// passes ping-pong (p, known); the last writes the state's p and p / rho^2.
template <typename S>
__global__ void k_step_closure_init(const S* __restrict__ rho, int64_t n, S thr, int* __restrict__ known,
                                    SimFlags* fl, int slot_in) {
  if (fl->stale[slot_in]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) known[i] = rho[i] < thr ? 0 : 1;
}

template <int D, typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k_step_closure(WcsphArgs<S> a, NB nbr, SimFlags* fl, int slot_in,
                                                         const S* __restrict__ p_src, const int* __restrict__ k_src,
                                                         S* __restrict__ p_dst, int* __restrict__ k_dst,
                                                         int literal, int last) {
  if (fl->stale[slot_in]) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  S pi = p_src[i];
  int ki = k_src[i];
  if (!ki) {
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
    Sum<S> num, den;
    nbr.template visit<D>(i, false, [&](int64_t j) {
      S df[D], r2 = S(0);
#pragma unroll
      for (int d = 0; d < D; ++d) { df[d] = (S)(xi[d] - a.pos[j * D + d]); r2 += df[d] * df[d]; }
      if (r2 > a.skip2) return;
      if (!literal && !k_src[j]) return;
      S w, g;
      cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
      num.add(w * (p_src[j] - (a.gravity * df[1]) * a.rho[j]));
      den.add(w);
    });
    if (den.get() > S(1e-30)) {
      const S v = num.get() / den.get();
      pi = v > S(0) ? v : S(0);
      ki = 1;
    }
  }
  p_dst[i] = pi;
  k_dst[i] = ki;
  if (last) {
    const S r = a.rho[i];
    if constexpr (sizeof(S) == 4) {
      // the closed pressure as stored (FP32 sums), its q in FP64
      const double q_d = (double)pi / ((double)r * (double)r);
      const S q = (S)q_d;
      a.pr2[i] = q;
      if (a.p_lo) a.p_lo[i] = S(0);
      if (a.aux) a.aux[i] = Vec4<S>{a.mass[i], q, a.vol[i], (S)(q_d - (double)q)};
    } else {
      a.pr2[i] = pi / (r * r);
    }
  }
}

template <typename T> TSPH_D T wave_max(T v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
template <typename T> TSPH_D T wave_min(T v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
TSPH_D double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One halo row as the force sweep stages it: the (m, p / rho^2, m / rho)
// record, position, and velocity in the sweep's precision (FP32 runs form
// the viscous velocity difference from the FP32 neighbour velocity in both
// the halo and the global path): 40 bytes in 2-D and 56 in 3-D for FP32, so
// a 3-D halo of up to 2,816 rows fits a CU's LDS.
template <int D, typename S> struct HaloRec {
  S q[4];                 // not Vec4: its 16-byte alignment would pad the record
  double x[D];
  S v[D];
};

template <int D, typename S, bool METERS, typename NB, bool HALO = false>
__global__ void __launch_bounds__(kBlock) k_step_force(WcsphArgs<S> a, NB nbr, S* __restrict__ acc_out,
                                                       StepArgs st, SimFlags* fl, int slot_in, double half_skin,
                                                       HaloView hv = {}, NbHalo hl = {}) {
  const int64_t nb = gridDim.x;
  const int64_t tile = xcd_block(blockIdx.x, nb);
  const int64_t i = tile * blockDim.x + threadIdx.x;
  if (fl->stale[slot_in]) {
    // binning no longer valid: carry the state through unchanged, keep the flag up
    if (i < a.n) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        st.pos_out[i * D + d] = a.pos[i * D + d];
        st.vel_out[i * D + d] = a.vel[i * D + d];
      }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) fl->stale[slot_in ^ 1] = 1;
    return;
  }
  int H = -1, Fh = 0;
  extern __shared__ __align__(16) unsigned char halo_lds[];
  HaloRec<D, S>* hrec = reinterpret_cast<HaloRec<D, S>*>(halo_lds);
  if constexpr (HALO) {
    if (tile < hv.ntiles) { const int2 c = hv.cnt[tile]; H = c.x; Fh = c.y; }
    if (H > hv.max_rows) H = -1;                     // more rows than this launch staged: global path
    if (H >= 0) {
      // every load of the staging is issued before the first store (up to
      // halo_cap / kBlock rows per thread)
      const uint32_t* hr = hv.rows + tile * halo_cap<D>();
      constexpr int QS = halo_cap<D>() / kBlock;
      double rx[QS][D];
      S rv[QS][D];
      S r0[QS], r1[QS], r2[QS], r3[QS];
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s = threadIdx.x + q * kBlock;
        const uint32_t j = hr[s < H ? s : 0];
#pragma unroll
        for (int d = 0; d < D; ++d) { rx[q][d] = a.pos[(int64_t)j * D + d]; rv[q][d] = (S)a.vel[(int64_t)j * D + d]; }
        if constexpr (sizeof(S) == 4) {
          const Vec4<S> w = a.aux[j];
          r0[q] = w.x; r1[q] = w.y; r2[q] = w.z; r3[q] = w.w;
        } else {
          r0[q] = a.mass[j]; r1[q] = a.pr2[j]; r2[q] = a.vol[j]; r3[q] = S(0);
        }
      }
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int s = threadIdx.x + q * kBlock;
        if (s < H) {
          HaloRec<D, S>& rr = hrec[s];
#pragma unroll
          for (int d = 0; d < D; ++d) { rr.x[d] = rx[q][d]; rr.v[d] = rv[q][d]; }
          rr.q[0] = r0[q]; rr.q[1] = r1[q]; rr.q[2] = r2[q]; rr.q[3] = r3[q];
        }
      }
      __syncthreads();
    }
  }
  double xi[D] = {}, vi[D] = {};
  S acc[D];
  double mi = 0.0, rhoi = 0.0, pi_ = 0.0;
  const bool live = i < a.n;
  if (live) {
#pragma unroll
    for (int d = 0; d < D; ++d) { xi[d] = a.pos[i * D + d]; vi[d] = a.vel[i * D + d]; }
    // FP32: the pressure terms as -2 q_i G_i (G_i from the density sweep, in
    // FP64) plus the contrast terms -m_j (q_j - q_i) grad W_ij, with q_j - q_i
    // formed from the two-part q of both rows (store_fluid_state); FP64: the
    // symmetric form -m_j (q_i + q_j) grad W_ij as it stands
    constexpr bool kSplit = sizeof(S) == 4;
    S pri, prl = S(0);
    if constexpr (kSplit) {
      const Vec4<S> ai = a.aux[i];
      pri = ai.y; prl = ai.w;
    } else {
      pri = a.pr2[i];
    }
    Sum<S> at[D];                                   // pressure + viscous terms
    // every list entry is evaluated, four at a time; an entry outside the
    // support has its gradient factor zeroed, which zeroes its pressure and
    // viscous terms (a frictionless wall pair's viscous term is zero through
    // its zero viscosity factor), so the sums hold the terms the branching
    // form adds.
    // fetch(u, xj, vj, q, wall) gives entry u of the chunk, from the halo or
    // from the arrays (32-bit byte offsets).
    // 2 nu for a fluid pair, 2 nu x friction for a wall pair (0 without friction)
    const S nu2w = a.nu2 * a.friction;
    auto chunk = [&](const bool* keep4, auto&& fetch) {
      S tt[D][4];
      // the four entries' loads first, then their arithmetic
      double pj4[4][D], vj4[4][D];
      Vec4<S> q4[4];                                // m, p / rho^2, m / rho
      bool wall4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) fetch(u, pj4[u], vj4[u], q4[u], wall4[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double* pj = pj4[u];
        const double* vj = vj4[u];
        const Vec4<S>& q = q4[u];
        S df[D], r2 = S(0);
#pragma unroll
        for (int d = 0; d < D; ++d) { df[d] = (S)(xi[d] - pj[d]); r2 += df[d] * df[d]; }
        const bool keep = keep4[u] && !(r2 > a.skip2);
        S g = cubic_g_r2<S>(r2, a.k);                // <= 0 inside the support
        g = keep ? g : S(0);
        S cp;
        if constexpr (kSplit) cp = -q.x * ((q.y - pri) + (q.w - prl));
        else cp = -q.x * (pri + q.y);
        const S cpg = cp * g;
        // Morris: -(2 nu) (m_j / rho_j) relu(-(x_ij . grad W)) / (r^2 + eta^2) v_ij,
        // with x_ij . grad W = g r^2 <= 0, so relu(-g r^2) = -g r^2
        const S nub = wall4[u] ? nu2w : a.nu2;
        const S cv = (nub * q.z) * quot<S>(g * r2, r2 + a.eta2);
#pragma unroll
        for (int d = 0; d < D; ++d) tt[d][u] = cpg * df[d] + cv * (S)(vi[d] - vj[d]);
      }
#pragma unroll
      for (int d = 0; d < D; ++d) at[d].add4(tt[d][0], tt[d][1], tt[d][2], tt[d][3]);
    };
    if (HALO && H >= 0) {
      hl.template visit_chunks<D, TSPH_HALO_CH>(i, [&](const int32_t* l4, const bool* keep4) {
        chunk(keep4, [&](int u, double (&pj)[D], double (&vj)[D], Vec4<S>& q, bool& wall) {
          const HaloRec<D, S>& rr = hrec[l4[u]];
#pragma unroll
          for (int d = 0; d < D; ++d) { pj[d] = rr.x[d]; vj[d] = rr.v[d]; }
          q = Vec4<S>{rr.q[0], rr.q[1], rr.q[2], rr.q[3]};
          wall = l4[u] >= Fh;
        });
      });
    } else {
      nbr.template visit_chunks<D, (D == 3 ? 2 : 1)>(i, true, [&](const int32_t* j4, const bool* keep4) {
        chunk(keep4, [&](int u, double (&pj)[D], double (&vj)[D], Vec4<S>& q, bool& wall) {
          const uint32_t j = (uint32_t)j4[u];
          const double* xp = at_bytes(a.pos, j * (uint32_t)(D * sizeof(double)));
          const double* vp = at_bytes(a.vel, j * (uint32_t)(D * sizeof(double)));
#pragma unroll
          for (int d = 0; d < D; ++d) { pj[d] = xp[d]; vj[d] = (double)(S)vp[d]; }
          if constexpr (sizeof(S) == 4) q = *at_bytes(a.aux, j * (uint32_t)sizeof(Vec4<S>));
          else q = Vec4<S>{a.mass[j], a.pr2[j], a.vol[j], S(0)};
          wall = j >= (uint32_t)a.n;
        });
      });
    }
    if constexpr (kSplit) {
      const double qi = (double)pri + (double)prl;
#pragma unroll
      for (int d = 0; d < D; ++d)
        acc[d] = (S)((hi(at[d]) - 2.0 * qi * (double)a.gsum[i * D + d]) + (d == 1 ? -(double)a.gravity : 0.0));
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d) acc[d] = at[d].get() + (d == 1 ? -a.gravity : S(0));
    }
    if (acc_out) {
#pragma unroll
      for (int d = 0; d < D; ++d) acc_out[i * D + d] = acc[d];
    }
    // the meters and the clamp impulse take the FP32 run's mass as its two parts
    mi = (double)a.mass[i] + (a.mass_lo ? (double)a.mass_lo[i] : 0.0);
    rhoi = (double)a.rho[i]; pi_ = (double)a.p[i];
  }
  // ---- integrate + clamp (droplet_impact.py:1745-1756, :1293-1312)
  const double dt = fl->step_dt;
  double disp = 0.0, spd2 = 0.0;
  unsigned long long hits = 0;
  if (live) {
    double xn[D], vn[D];
    int hit_any = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      vn[d] = vi[d] + dt * (double)acc[d];
      double x = xi[d] + dt * vn[d];
      const int hit = (x < st.lo[d] - st.tol) || (x > st.hi[d] + st.tol);
      hit_any |= hit;
      x = x < st.lo[d] ? st.lo[d] : (x > st.hi[d] ? st.hi[d] : x);
      if (hit) vn[d] = 0.0;
      xn[d] = x;
    }
    obstacle_clamp<D>(xn, vn, mi, st);
#pragma unroll
    for (int d = 0; d < D; ++d) { st.pos_out[i * D + d] = xn[d]; st.vel_out[i * D + d] = vn[d]; }
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) { const double df = xn[d] - st.anchor[i * D + d]; d2 += df * df; }
    disp = sqrt(d2);
#pragma unroll
    for (int d = 0; d < D; ++d) spd2 += vi[d] * vi[d];
    hits = hit_any;
  }
  // ---- per-step maxima (density, speed, displacement), one atomic per block
  const double wdisp = wave_max(disp);
  const double wspd2 = wave_max(spd2);
  const double wrho = wave_max(live ? rhoi : -1e300);
  unsigned long long whits = hits;
  for (int o = 32; o > 0; o >>= 1) whits += __shfl_xor(whits, o, 64);
  __shared__ double bmax[3][kBlock / 64];
  __shared__ unsigned long long bhits[kBlock / 64];
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    bmax[0][w] = wdisp; bmax[1][w] = wspd2; bmax[2][w] = wrho; bhits[w] = whits;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bd = bmax[0][0], bv = bmax[1][0], br = bmax[2][0];
    unsigned long long bh = bhits[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      bd = fmax(bd, bmax[0][w]); bv = fmax(bv, bmax[1][w]); br = fmax(br, bmax[2][w]); bh += bhits[w];
    }
    const double bspd = sqrt(bv);
    if (st.block_max) {
      // plain stores: thousands of blocks contending for one cache line
      // with atomics cost more than the sweep itself
      unsigned long long* o = st.block_max + 4 * (int64_t)blockIdx.x;
      o[0] = ord_enc(bd); o[1] = ord_enc(bspd); o[2] = ord_enc(br); o[3] = bh;
    } else {
      atomicMax(&fl->maxes[MM_DISP], ord_enc(bd));
      atomicMax(&fl->maxes[MM_PEAK_SPEED], ord_enc(bspd));
      atomicMax(&fl->maxes[MM_PEAK_RHO], ord_enc(br));
      if (bh) atomicAdd(&fl->escaped, bh);
    }
    if (METERS) {
      atomicMax(&fl->maxes[MM_SPEED], ord_enc(bspd));
      atomicMax(&fl->maxes[MM_RHO_MAX], ord_enc(br));
    }
    if (bd > half_skin) fl->stale[slot_in ^ 1] = 1;
  }
  if (METERS) {
    // sums of the pre-integration state (the rows of droplet_impact.py:1709-1733)
    double accn2 = 0.0;
    if (live) {
#pragma unroll
      for (int d = 0; d < D; ++d) accn2 += (double)acc[d] * (double)acc[d];
    }
    double s[MS_NSUM];
    s[MS_MASS] = mi;
    s[MS_MX] = mi * vi[0];
    s[MS_MY] = D > 1 ? mi * vi[1 % D] : 0.0;
    s[MS_MZ] = D > 2 ? mi * vi[2 % D] : 0.0;
    s[MS_KE] = live ? 0.5 * (mi * spd2) : 0.0;
    s[MS_RHO] = live ? rhoi : 0.0;
    s[MS_PE] = live ? mi * st.gravity * xi[1 % D] : 0.0;
    __shared__ double red[MS_NSUM][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MS_NSUM; ++k) {
      const double v = wave_sum(s[k]);
      if (lane == 0) red[k][wave] = v;
    }
    const double big = 1e300;
    // row extrema, ordered so that all are maxima: p, |acc|, y, x, -rho, -p, -y
    double ext[7] = {wave_max(live ? pi_ : -big), wave_max(sqrt(accn2)), wave_max(live ? xi[1 % D] : -big),
                     wave_max(live ? xi[0] : -big), -wave_min(live ? rhoi : big), -wave_min(live ? pi_ : big),
                     -wave_min(live ? xi[1 % D] : big)};
    __shared__ double bext[7][kBlock / 64];
    if (lane == 0)
      for (int k = 0; k < 7; ++k) bext[k][wave] = ext[k];
    __syncthreads();
    if (threadIdx.x < MS_NSUM) {
      double v = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[threadIdx.x][w];
      st.block_sums[(int64_t)blockIdx.x * MS_NSUM + threadIdx.x] = v;
    }
    if (threadIdx.x == 0) {                 // one atomic per block and extremum
      for (int k = 0; k < 7; ++k)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) ext[k] = fmax(ext[k], bext[k][w]);
      atomicMax(&fl->maxes[MM_P_MAX], ord_enc(ext[0]));
      atomicMax(&fl->maxes[MM_ACC], ord_enc(ext[1]));
      atomicMax(&fl->maxes[MM_Y_MAX], ord_enc(ext[2]));
      atomicMax(&fl->maxes[MM_X_MAX], ord_enc(ext[3]));
      atomicMin(&fl->mins[MN_RHO_MIN], ord_enc(-ext[4]));
      atomicMin(&fl->mins[MN_P_MIN], ord_enc(-ext[5]));
      atomicMin(&fl->mins[MN_Y_MIN], ord_enc(-ext[6]));
    }
  }
}

// fixed-order reduction of the per-block partial sums
__global__ void k_reduce_sums(const double* __restrict__ part, int64_t nblocks, double* __restrict__ out,
                              SimFlags* fl, int slot_in) {
  if (fl->stale[slot_in]) return;
  __shared__ double red[MS_NSUM][256];
  for (int k = 0; k < MS_NSUM; ++k) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[b * MS_NSUM + k];
    red[k][threadIdx.x] = v;
  }
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride)
      for (int k = 0; k < MS_NSUM; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x < MS_NSUM) out[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_reset_meters(SimFlags* fl, int slot_in) {
  if (fl->stale[slot_in]) return;
  const int t = threadIdx.x;
  if (t < MM_NMAX && t != MM_DISP && t != MM_PEAK_SPEED && t != MM_PEAK_RHO) fl->maxes[t] = 0;
  if (t < MN_NMIN) fl->mins[t] = ~0ull;
}

// ------------------------------------------------------------------ rebuild helpers
// Offset of the low parts of the masses in the mass buffer (FP32 runs), else 0.
inline int64_t mass_lo_offset(const tsph_sim* sim) { return sim->prec == TSPH_F64 ? 0 : sim->nt; }

template <int D, typename S>
__global__ void k_permute(const int64_t* __restrict__ new2old, int64_t nt, const double* __restrict__ pos,
                          const double* __restrict__ vel, const S* __restrict__ mass,
                          const int64_t* __restrict__ orig, const int64_t* __restrict__ cells_old,
                          double* __restrict__ pos2, double* __restrict__ vel2, S* __restrict__ mass2,
                          int64_t* __restrict__ orig2, int64_t* __restrict__ cells_new, int64_t lo_off) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const int64_t o = new2old[k];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    pos2[k * D + d] = pos[o * D + d];
    vel2[k * D + d] = vel[o * D + d];
    cells_new[k * D + d] = cells_old[o * D + d];
  }
  mass2[k] = mass[o];
  if (lo_off) mass2[lo_off + k] = mass[lo_off + o];    // the low parts (FP32 runs)
  orig2[k] = orig[o];
}

// ------------------------------------------------------------------ scatter back
template <int D, typename S>
__global__ void k_unpermute(const int64_t* __restrict__ orig, int64_t n, const double* __restrict__ pos,
                            const double* __restrict__ vel, const S* __restrict__ rho,
                            const S* __restrict__ p, const S* __restrict__ acc, double* __restrict__ pos_o,
                            double* __restrict__ vel_o, S* __restrict__ rho_o, S* __restrict__ p_o,
                            S* __restrict__ acc_o) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t o = orig[k];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (pos_o) pos_o[o * D + d] = pos[k * D + d];
    if (vel_o) vel_o[o * D + d] = vel[k * D + d];
    if (acc_o) acc_o[o * D + d] = acc[k * D + d];
  }
  if (rho_o) rho_o[o] = rho[k];
  if (p_o) p_o[o] = p[k];
}

// unordered pairs with |x_i - x_j| <= radius: fluid-fluid pairs are listed
// from both ends (counted at the lower row), fluid-wall pairs once
template <int D>
__global__ void k_count_close(const double* __restrict__ pos, int64_t n, NbCells nbr, double r2max,
                              unsigned long long* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  if (i < n) {
    nbr.template visit<D>(i, true, [&](int64_t j) {
      if (j < n && j < i) return;
      double r2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) { const double df = pos[i * D + d] - pos[j * D + d]; r2 += df * df; }
      if (r2 <= r2max) ++c;
    });
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// row k's mass from the caller's FP64 masses: FP32 runs as hi + lo (lo at
// lo_off), FP64 runs as given
template <typename S>
__global__ void k_set_masses(const double* __restrict__ m, const int64_t* __restrict__ orig, int64_t nt,
                             S* __restrict__ mass, int64_t lo_off) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const double v = m[orig[k]];
  const S h = (S)v;
  mass[k] = h;
  if (lo_off) mass[lo_off + k] = (S)(v - (double)h);
}

__global__ void k_iota(int64_t* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

}  // namespace tsph

using namespace tsph;

namespace tsph {

// Dense per-cell row ranges of the rows in their current order (fluid and
// wall rows each grouped by cell; cells_new holds every row's cell) and the
// per-row lists at the build radius rc in fixed slots, walked from them.
// The dense cell table of rows sorted by cell (fluid rows, then walls):
// row-major strides (last axis fastest) unless given.
template <int D>
int build_cell_table(tsph_sim* sim, const int64_t* ext, const int64_t* stride, hipStream_t stream) {
  const int64_t n = sim->n, nt = sim->nt;
  double dense = 1.0;
  for (int d = 0; d < D; ++d) dense *= (double)ext[d];
  TSPH_REQUIRE(dense <= (double)kMaxDenseCells, TSPH_ERR_ARG,
               "cell grid of %.3g cells exceeds the dense range table (%lld cells): the particle "
               "cloud is too sparse for its cutoff", dense, (long long)kMaxDenseCells);
  sim->ncells_dense = (int64_t)dense;
  int64_t rm[3] = {0, 0, 0};
  rm[D - 1] = 1;
  for (int d = D - 2; d >= 0; --d) rm[d] = rm[d + 1] * ext[d + 1];
  for (int d = 0; d < 3; ++d) {
    sim->cell_ext[d] = d < D ? ext[d] : 1;
    sim->cell_stride[d] = d < D ? (stride ? stride[d] : rm[d]) : 0;
  }
  TSPH_TRY(sim->cell_R.ensure(sizeof(int4) * (size_t)sim->ncells_dense));
  k_cell_table_clear<<<grid_for(sim->ncells_dense), kBlock, 0, stream>>>(sim->cell_R.as<int4>(), sim->ncells_dense);
  k_cell_table<D><<<grid_for(nt), kBlock, 0, stream>>>(sim->cells_new.as<int64_t>(), nt, n, sim->cell_stride[0],
                                                       sim->cell_stride[1], sim->cell_stride[2],
                                                       sim->cell_R.as<int4>());
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

template <int D>
int build_sim_slots(tsph_sim* sim, const double* npos, double rc, const int64_t* ext, const int64_t* stride,
                    hipStream_t stream) {
  const int64_t n = sim->n, nt = sim->nt;
  sim->halo_ok = false;                   // halos name the rows of the lists they were built from
  TSPH_TRY(build_cell_table<D>(sim, ext, stride, stream));
  // every row lies within ext[d] cells of edge rc of row 0
  double extent = 0.0;
  for (int d = 0; d < D; ++d) extent = fmax(extent, (double)sim->cell_ext[d] * rc);
  return build_slot_lists<D, false>(sim_cells(sim), npos, n, nt, rc * rc, nullptr, sim->slot_cnt, sim->slot_idx,
                                    sim->slot_max, sim->slot_K, nullptr, stream, &sim->slot_posf, extent);
}
template int build_sim_slots<2>(tsph_sim*, const double*, double, const int64_t*, const int64_t*, hipStream_t);

// The halo layout's lists (tiles == 2): per tile of kHaloRows fluid rows
// one workgroup builds the tile's global and halo lists and its halo
// (k_halo_lists); the wall rows' lists come from the cell walk; the reduce
// kernel writes the build's report into pinned memory and marks the step
// flags stale when a list did not fit its slots or a row was binned from
// outside the grid.  lagged: one pass at the current width, no wait; the
// host reads the report when the steps on this build come back (a stale
// report makes it build again, wider).  Otherwise passes are repeated, wider,
// until every list fits (the first build counts first), and the stream is
// synchronised.  The sweeps stage halos of up to halo_max rows (a tile with a
// larger halo reads the global arrays): the exact largest halo after a
// synchronised build, the last report's plus a margin after a lagged one,
// within halo_lds_budget<D>() bytes of staged rows (rec_bytes each): in 2-D the
// default dynamic LDS limit of a launch (four force workgroups per CU), in
// 3-D the whole 160 KiB of a CU, which the 3-D halos (1,500 - 2,000 rows)
// need; the sweeps raise their launch limit to it (allow_halo_lds).
// (the sweeps' own static LDS, a few hundred bytes of reductions, takes
// part of the last 4 KiB)
template <int D> constexpr size_t halo_lds_budget() { return D == 2 ? 60 * 1024 : 156 * 1024; }
constexpr int kHaloMargin = 64;

// Raises a halo sweep's dynamic LDS limit to halo_lds_budget<3>() once per
// kernel (its static LDS and the budget must fit the CU's 160 KiB).
template <auto KERNEL>
int allow_halo_lds() {
  static const int status = [] {
    const void* fn = reinterpret_cast<const void*>(KERNEL);
    hipFuncAttributes attr{};
    hipError_t e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) { set_error("hipFuncGetAttributes: %s", hipGetErrorString(e)); return TSPH_ERR_HIP; }
    if (attr.sharedSizeBytes + halo_lds_budget<3>() > 160 * 1024) {
      set_error("halo sweep: %zu bytes of static LDS leave no room for %zu staged", (size_t)attr.sharedSizeBytes,
                halo_lds_budget<3>());
      return TSPH_ERR_HIP;
    }
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)halo_lds_budget<3>());
    if (e != hipSuccess) { set_error("hipFuncSetAttribute: %s", hipGetErrorString(e)); return TSPH_ERR_HIP; }
    return TSPH_OK;
  }();
  return status;
}

template <int D>
int build_halo_lists(tsph_sim* sim, const double* npos, double rc, size_t rec_bytes, bool lagged,
                     hipStream_t stream) {
  const int64_t n = sim->n, nt = sim->nt;
  sim->halo_ok = false;
  TSPH_TRY(build_cell_table<D>(sim, sim->bgrid.ext, nullptr, stream));
  const int64_t ntiles = (n + kHaloRows - 1) / kHaloRows;
  sim->halo_tiles = ntiles;
  const int64_t groups = (nt + 63) / 64;
  constexpr int kCap = halo_cap<D>();
  const int cap = (int)std::min<size_t>((size_t)kCap, halo_lds_budget<D>() / rec_bytes);
  const double rc2 = rc * rc;
  TSPH_TRY(sim->slot_cnt.ensure(sizeof(int32_t) * (size_t)(nt > 0 ? nt : 1)));
  TSPH_TRY(sim->slot_max.ensure(sizeof(unsigned int)));
  TSPH_TRY(sim->halo_rows.ensure(sizeof(uint32_t) * (size_t)(ntiles > 0 ? ntiles : 1) * kCap));
  TSPH_TRY(sim->halo_cnt.ensure(sizeof(int2) * (size_t)(ntiles > 0 ? ntiles : 1)));
  TSPH_TRY(sim->halo_hmax.ensure(sizeof(unsigned int) * (2 * (size_t)ntiles + 2)));
  unsigned int* blk = sim->halo_hmax.as<unsigned int>();
  SimFlags* fl = sim->flags.as<SimFlags>();
  int64_t& K = sim->slot_K;
  const bool wait = !lagged || K == 0;
  const unsigned int prev_hmax = sim->build_stats->hmax;
  for (int pass = 0; pass < 4; ++pass) {
    int32_t* gidx = nullptr;
    uint16_t* hidx = nullptr;
    if (K > 0) {
      TSPH_TRY(sim->slot_idx.ensure(sizeof(int32_t) * (size_t)groups * 64 * (size_t)K));
      TSPH_TRY(sim->halo_idx.ensure(sizeof(uint16_t) * (size_t)groups * 64 * (size_t)K + 64));
      gidx = sim->slot_idx.as<int32_t>();
      hidx = sim->halo_idx.as<uint16_t>();
    }
    TSPH_CHECK_HIP(hipMemsetAsync(&fl->stale[0], 0, sizeof(unsigned long long) * 2, stream));
    TSPH_CHECK_HIP(hipMemsetAsync(sim->slot_max.ptr, 0, sizeof(unsigned int), stream));
    if (ntiles > 0) {
      if (K > 0)
        k_halo_lists<D, true><<<(unsigned)ntiles, kHaloRows, 0, stream>>>(
            sim_cells(sim), npos, n, rc2, cap, sim->slot_cnt.as<int32_t>(), gidx, hidx, K,
            sim->halo_rows.as<uint32_t>(), sim->halo_cnt.as<int2>(), blk);
      else
        k_halo_lists<D, false><<<(unsigned)ntiles, kHaloRows, 0, stream>>>(
            sim_cells(sim), npos, n, rc2, cap, sim->slot_cnt.as<int32_t>(), gidx, hidx, K,
            sim->halo_rows.as<uint32_t>(), sim->halo_cnt.as<int2>(), blk);
    }
    if (nt > n) {
      const unsigned g = grid_xcd(nt - n);
      if (K > 0)
        k_slot_lists<D, true, false, false><<<g, kBlock, 0, stream>>>(
            sim_cells(sim), npos, nullptr, n, n, nt, rc2, 0.f, 0.f, 0.0, 0.0, 0.0, sim->slot_cnt.as<int32_t>(),
            gidx, K, sim->slot_max.as<unsigned int>(), nullptr);
      else
        k_slot_lists<D, false, false, false><<<g, kBlock, 0, stream>>>(
            sim_cells(sim), npos, nullptr, n, n, nt, rc2, 0.f, 0.f, 0.0, 0.0, 0.0, sim->slot_cnt.as<int32_t>(),
            gidx, K, sim->slot_max.as<unsigned int>(), nullptr);
    }
    k_halo_lists_reduce<<<1, kHaloRows, 0, stream>>>(blk, ntiles, sim->slot_max.as<unsigned int>(),
                                                     sim->bin_flag.as<unsigned int>(), K, &fl->stale[0],
                                                     sim->build_stats);
    TSPH_CHECK_LAUNCH();
    if (!wait) {
      // prev_hmax: the largest halo of the last build that reported
      sim->halo_max = std::min(cap, (int)prev_hmax + kHaloMargin);
      sim->halo_ok = true;
      return TSPH_OK;
    }
    TSPH_CHECK_HIP(hipStreamSynchronize(stream));
    const BuildStats st = *sim->build_stats;
    if (st.outside) return TSPH_OK;               // the caller regrids (halo not in place)
    if (K > 0 && (int64_t)st.kmax <= K) {
      sim->halo_max = (int)st.hmax;
      sim->halo_ok = true;
      return TSPH_OK;
    }
    K = ((int64_t)st.kmax + st.kmax / 8 + 4 + 3) / 4 * 4;
  }
  TSPH_REQUIRE(false, TSPH_ERR_HIP, "neighbour lists did not fit %lld slots per row", (long long)K);
  return TSPH_OK;
}

// The 3-D halo layout's halos, over the lists the cell walk just built
// (build_sim_slots, synchronised): per tile the halo and the halo lists of
// the global lists (k_halo_lists in MAP form); synchronises for the largest
// halo.
template <int D>
int build_halo_map(tsph_sim* sim, size_t rec_bytes, hipStream_t stream) {
  const int64_t n = sim->n, nt = sim->nt;
  sim->halo_ok = false;
  const int64_t ntiles = (n + kHaloRows - 1) / kHaloRows;
  sim->halo_tiles = ntiles;
  const int64_t groups = (nt + 63) / 64;
  constexpr int kCap = halo_cap<D>();
  const int cap = (int)std::min<size_t>((size_t)kCap, halo_lds_budget<D>() / rec_bytes);
  const int64_t K = sim->slot_K;
  TSPH_TRY(sim->halo_rows.ensure(sizeof(uint32_t) * (size_t)(ntiles > 0 ? ntiles : 1) * kCap));
  TSPH_TRY(sim->halo_cnt.ensure(sizeof(int2) * (size_t)(ntiles > 0 ? ntiles : 1)));
  TSPH_TRY(sim->halo_hmax.ensure(sizeof(unsigned int) * (2 * (size_t)ntiles + 2)));
  TSPH_TRY(sim->halo_idx.ensure(sizeof(uint16_t) * (size_t)groups * 64 * (size_t)K + 64));
  unsigned int* blk = sim->halo_hmax.as<unsigned int>();
  SimFlags* fl = sim->flags.as<SimFlags>();
  if (ntiles > 0)
    k_halo_lists<D, true, true><<<(unsigned)ntiles, kHaloRows, 0, stream>>>(
        sim_cells(sim), sim->pos[sim->cur].as<double>(), n, 0.0, cap, sim->slot_cnt.as<int32_t>(),
        sim->slot_idx.as<int32_t>(), sim->halo_idx.as<uint16_t>(), K, sim->halo_rows.as<uint32_t>(),
        sim->halo_cnt.as<int2>(), blk);
  k_halo_lists_reduce<<<1, kHaloRows, 0, stream>>>(blk, ntiles, sim->slot_max.as<unsigned int>(),
                                                   sim->bin_flag.as<unsigned int>(), K, &fl->stale[0],
                                                   sim->build_stats);
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  sim->halo_max = (int)sim->build_stats->hmax;
  sim->halo_ok = true;
  return TSPH_OK;
}

static inline HaloView sim_halo(const tsph_sim* sim) {
  HaloView h{};
  h.rows = sim->halo_rows.as<uint32_t>();
  h.cnt = sim->halo_cnt.as<int2>();
  h.idx = sim->halo_idx.as<uint16_t>();
  h.ntiles = sim->halo_tiles;
  h.max_rows = sim->halo_max;
  return h;
}
static inline NbHalo sim_halo_lists(const tsph_sim* sim) {
  NbHalo l{};
  l.idx = sim->halo_idx.as<uint16_t>();
  l.cnt = sim->slot_cnt.as<int32_t>();
  l.K = sim->slot_K;
  return l;
}

// Stable counting sort of rows [r0, r1) by their bin keys (k_bin has counted
// them into bin_cnt): new2old[r0, r1) lists the rows in key order, ties in
// row order.  bin_cnt is all zero again afterwards.
int bin_sort(tsph_sim* sim, int64_t r0, int64_t r1, hipStream_t stream) {
  if (r1 <= r0) return TSPH_OK;
  const int64_t M = sim->bgrid.M;
  const int64_t nparts = (M + kScanTile - 1) / kScanTile;
  int32_t* cnt = sim->bin_cnt.as<int32_t>();
  int32_t* start = sim->bin_start.as<int32_t>();
  int32_t* part = sim->bin_part.as<int32_t>();
  k_scan_sums<<<(unsigned)nparts, kBlock, 0, stream>>>(cnt, M, part);
  k_scan_parts<<<1, kBlock, 0, stream>>>(part, nparts);
  k_scan_apply<<<(unsigned)nparts, kBlock, 0, stream>>>(cnt, M, part, (int32_t)r0, start);
  const unsigned g = grid_for(r1 - r0);
  k_bin_scatter<<<g, kBlock, 0, stream>>>(sim->bin_keys.as<int32_t>(), r0, r1, start, cnt, sim->bin_tmp.as<int32_t>());
  k_bin_rank<<<g, kBlock, 0, stream>>>(sim->bin_tmp.as<int32_t>(), sim->bin_keys.as<int32_t>(), r0, r1, start,
                                       sim->new2old.as<int64_t>());
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// The binning grid: cells of edge rc over the clamp box and every row now
// (synchronises).  Sizes the sort's buffers.
template <int D>
int bin_regrid(tsph_sim* sim, double rc, hipStream_t stream) {
  const int64_t nt = sim->nt;
  const unsigned g = grid_for(nt);
  TSPH_TRY(sim->bin_bounds.ensure(sizeof(long long) * (6 * (size_t)g + 6)));
  long long* blk = sim->bin_bounds.as<long long>();
  k_bin_bounds<D><<<g, kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), nt, rc, blk);
  k_bin_bounds_reduce<<<1, kBlock, 0, stream>>>(blk, (int64_t)g, blk + 6 * (size_t)g);
  TSPH_CHECK_LAUNCH();
  long long mm[6];
  TSPH_CHECK_HIP(hipMemcpyAsync(mm, blk + 6 * (size_t)g, sizeof(mm), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  BinGrid& bg = sim->bgrid;
  bg = BinGrid{};
  bg.rc = rc;
  double cells = 1.0;
  int64_t blocks = 1;
  for (int d = 0; d < 3; ++d) { bg.cmin[d] = 0; bg.ext[d] = 1; bg.nb[d] = 1; }
  for (int d = 0; d < D; ++d) {
    const long long blo = (long long)floor(sim->prm.lo[d] / rc), bhi = (long long)floor(sim->prm.hi[d] / rc);
    const long long lo = mm[d] < blo ? mm[d] : blo, hi = mm[3 + d] > bhi ? mm[3 + d] : bhi;
    bg.cmin[d] = lo;
    bg.ext[d] = hi - lo + 1;
    bg.nb[d] = (bg.ext[d] + BinOrder<D>::B - 1) / BinOrder<D>::B;
    cells *= (double)bg.ext[d];
    blocks *= bg.nb[d];
  }
  TSPH_REQUIRE(cells <= (double)kMaxDenseCells && (double)blocks * BinOrder<D>::PER_BLOCK < 2.0e9, TSPH_ERR_ARG,
               "cell grid of %.3g cells exceeds the dense range table (%lld cells): the particle "
               "cloud or the clamp box is too sparse for its cutoff", cells, (long long)kMaxDenseCells);
  bg.M = blocks * BinOrder<D>::PER_BLOCK;
  // the counts return to zero after every sort; a regrid starts them there
  TSPH_TRY(sim->bin_cnt.ensure(sizeof(int32_t) * (size_t)(bg.M + 1)));
  TSPH_CHECK_HIP(hipMemsetAsync(sim->bin_cnt.ptr, 0, sizeof(int32_t) * (size_t)(bg.M + 1), stream));
  TSPH_TRY(sim->bin_start.ensure(sizeof(int32_t) * (size_t)(bg.M + 1)));
  TSPH_TRY(sim->bin_part.ensure(sizeof(int32_t) * (size_t)((bg.M + kScanTile - 1) / kScanTile)));
  TSPH_TRY(sim->bin_keys.ensure(sizeof(int32_t) * (size_t)nt));
  TSPH_TRY(sim->bin_tmp.ensure(sizeof(int32_t) * (size_t)nt));
  TSPH_TRY(sim->bin_cells.ensure(sizeof(int64_t) * (size_t)nt * D));
  TSPH_TRY(sim->bin_flag.ensure(sizeof(unsigned int)));
  sim->bgrid_ok = true;
  return TSPH_OK;
}

// New row order: the fluid rows by key (a full build also regrids and sorts
// the walls, which keep their rows until the next full build).  Fills
// new2old and bin_cells (cell of every row in the old order).
template <int D>
int bin_rows(tsph_sim* sim, bool full, double rc, hipStream_t stream) {
  const int64_t n = sim->n, nt = sim->nt;
  if (full) TSPH_TRY(bin_regrid<D>(sim, rc, stream));
  const double* pos = sim->pos[sim->cur].as<double>();
  TSPH_CHECK_HIP(hipMemsetAsync(sim->bin_flag.ptr, 0, sizeof(unsigned int), stream));
  auto part = [&](int64_t r0, int64_t r1) -> int {
    if (r1 <= r0) return TSPH_OK;
    k_bin<D><<<grid_for(r1 - r0), kBlock, 0, stream>>>(pos, r0, r1, sim->bgrid, sim->bin_cells.as<int64_t>(),
                                                       sim->bin_keys.as<int32_t>(), sim->bin_cnt.as<int32_t>(),
                                                       sim->bin_flag.as<unsigned int>());
    TSPH_CHECK_LAUNCH();
    return bin_sort(sim, r0, r1, stream);
  };
  TSPH_TRY(part(0, n));
  if (full) TSPH_TRY(part(n, nt));
  TSPH_CHECK_HIP(hipMemcpyAsync(sim->bin_outside, sim->bin_flag.ptr, sizeof(unsigned int),
                                hipMemcpyDeviceToHost, stream));
  return TSPH_OK;
}

// Re-bins the rows and builds the lists.  lagged (the stepping loop): the
// 2-D halo build does not wait for its report; the report of the previous
// build, complete by now, widens the slots or regrids when that build's lists
// did not fit or a row was outside the grid, and such a repeat of a build on
// unchanged positions is not counted.
template <int D, typename S>
int sim_rebuild(tsph_sim* sim, hipStream_t stream, bool lagged = false) {
  const int64_t n = sim->n, nt = sim->nt;
  const double rc = 2.0 * sim->prm.op.h + sim->prm.skin;
  const bool halo_lists = sim->tiles == 2;
  if (sim->slot_K_next >= 0) { sim->slot_K = sim->slot_K_next; sim->slot_K_next = -1; }
  const BuildStats prev = *sim->build_stats;
  const bool redo = lagged && halo_lists && sim->built && prev.K > 0 && (prev.kmax > prev.K || prev.outside);
  if (redo && (int64_t)prev.kmax > sim->slot_K) sim->slot_K = ((int64_t)prev.kmax + prev.kmax / 8 + 4 + 3) / 4 * 4;
  // rows in blocked Morton order of their cells: a run of consecutive
  // targets and all their neighbours sit in a compact block of memory
  bool full = !sim->bgrid_ok || sim->bgrid.rc != rc || (redo && prev.outside);
  for (int attempt = 0; attempt < 2; ++attempt) {
    TSPH_TRY(bin_rows<D>(sim, full, rc, stream));
    const int64_t rows = full ? nt : n;        // a full build moves the walls too
    const int other = sim->cur ^ 1;
    k_permute<D, S><<<grid_for(rows), kBlock, 0, stream>>>(
        sim->new2old.as<int64_t>(), rows, sim->pos[sim->cur].as<double>(), sim->vel[sim->cur].as<double>(),
        sim->mass.as<S>(), sim->orig.as<int64_t>(), sim->bin_cells.as<int64_t>(), sim->pos[other].as<double>(),
        sim->vel[other].as<double>(), sim->mass2.as<S>(), sim->orig2.as<int64_t>(), sim->cells_new.as<int64_t>(),
        mass_lo_offset(sim));
    TSPH_CHECK_LAUNCH();
    std::swap(sim->mass, sim->mass2);
    std::swap(sim->orig, sim->orig2);
    sim->cur = other;
    if (full && nt > n) {
      // both ping-pong sets carry the (static) wall rows
      const int o = sim->cur ^ 1;
      TSPH_CHECK_HIP(hipMemcpyAsync(sim->pos[o].as<double>() + n * D, sim->pos[sim->cur].as<double>() + n * D,
                                    sizeof(double) * (nt - n) * D, hipMemcpyDeviceToDevice, stream));
      TSPH_CHECK_HIP(hipMemcpyAsync(sim->vel[o].as<double>() + n * D, sim->vel[sim->cur].as<double>() + n * D,
                                    sizeof(double) * (nt - n) * D, hipMemcpyDeviceToDevice, stream));
      TSPH_CHECK_HIP(hipMemcpyAsync(sim->mass2.as<S>() + n, sim->mass.as<S>() + n, sizeof(S) * (nt - n),
                                    hipMemcpyDeviceToDevice, stream));
      if (const int64_t lo = mass_lo_offset(sim))
        TSPH_CHECK_HIP(hipMemcpyAsync(sim->mass2.as<S>() + lo + n, sim->mass.as<S>() + lo + n,
                                      sizeof(S) * (nt - n), hipMemcpyDeviceToDevice, stream));
      TSPH_CHECK_HIP(hipMemcpyAsync(sim->orig2.as<int64_t>() + n, sim->orig.as<int64_t>() + n,
                                    sizeof(int64_t) * (nt - n), hipMemcpyDeviceToDevice, stream));
    }
    const double* npos = sim->pos[sim->cur].as<double>();
    if (halo_lists && D == 2) {
      // clears the step flags; a lagged build may mark them stale again
      TSPH_TRY(build_halo_lists<D>(sim, npos, rc, sizeof(HaloRec<D, S>), lagged, stream));
      if (lagged && sim->slot_K > 0 && sim->halo_ok) break;
      if (!sim->build_stats->outside) break;
    } else {
      TSPH_TRY((build_sim_slots<D>(sim, npos, rc, sim->bgrid.ext, nullptr, stream)));  // synchronises
      SimFlags* fl = sim->flags.as<SimFlags>();
      TSPH_CHECK_HIP(hipMemsetAsync(&fl->stale[0], 0, sizeof(unsigned long long) * 2, stream));
      if (!*sim->bin_outside) {
        if (halo_lists) TSPH_TRY(build_halo_map<D>(sim, sizeof(HaloRec<D, S>), stream));
        break;
      }
    }
    // a fluid row left the grid (a correction moved it past the clamp box):
    // regrid and bin every row again
    TSPH_REQUIRE(!full, TSPH_ERR_HIP, "binning grid does not cover the rows after a full build");
    full = true;
  }
  const double* npos = sim->pos[sim->cur].as<double>();
  TSPH_CHECK_HIP(hipMemcpyAsync(sim->anchor.ptr, npos, sizeof(double) * n * D, hipMemcpyDeviceToDevice, stream));
  if (!redo) sim->rebuilds++;
  sim->built = true;
  return TSPH_OK;
}

template <int D, typename S>
WcsphArgs<S> sim_args(tsph_sim* sim) {
  WcsphArgs<S> a = make_wcsph_args<S>(sim->prm.op, D);
  a.pos = sim->pos[sim->cur].as<double>();
  a.vel = sim->vel[sim->cur].as<double>();
  a.mass = sim->mass.as<S>();
  a.mass_lo = mass_lo_offset(sim) ? sim->mass.as<S>() + mass_lo_offset(sim) : nullptr;
  a.wall_accel = nullptr;
  a.rho = sim->rho.as<S>(); a.p = sim->p.as<S>(); a.pr2 = sim->pr2.as<S>(); a.vol = sim->vol.as<S>();
  // FP32: one 16-byte record per neighbour; FP64 reads the three arrays
  a.aux = sizeof(S) == 4 ? sim->aux.as<Vec4<S>>() : nullptr;
  a.p_lo = sizeof(S) == 4 ? sim->p_lo.as<S>() : nullptr;
  a.gsum = sizeof(S) == 4 ? sim->gsum.as<S>() : nullptr;
  return a;
}

// After the fluid density and EOS: the optional free-surface pressure closure,
// the wall sweep, and the optional boundary-consistent density (fluid
// correction, then the wall sweep again), in the reference's order.
template <int D, typename S>
int queue_wall_stage(tsph_sim* sim, const WcsphArgs<S>& a, const NbSlots& nbr, SimFlags* fl, int slot,
                     hipStream_t stream) {
  const tsph_wcsph_params& op = sim->prm.op;
  const int64_t n = sim->n, nw = sim->nw;
  if (op.surface_pressure_closure && n > 0) {
    const int layers = op.surface_pressure_donors_all ? 1
                       : (op.surface_pressure_layers > 1 ? op.surface_pressure_layers : 1);
    const size_t sb = (sizeof(S) * (size_t)(n + 1) + 255) / 256 * 256;
    const size_t kb = (sizeof(int) * (size_t)(n + 1) + 255) / 256 * 256;
    char* base = sim->closure.as<char>();
    S* pbuf[2] = {(S*)base, (S*)(base + sb)};
    int* kbuf[2] = {(int*)(base + 2 * sb), (int*)(base + 2 * sb + kb)};
    TSPH_CHECK_HIP(hipMemcpyAsync(pbuf[0], a.p, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
    k_step_closure_init<S><<<grid_for(n), kBlock, 0, stream>>>(a.rho, n, S(op.surface_pressure_threshold * op.rho0),
                                                               kbuf[0], fl, slot);
    for (int l = 0; l < layers; ++l) {
      const int src = l & 1, last = l == layers - 1;
      k_step_closure<D, S, NbSlots><<<grid_for(n), kBlock, 0, stream>>>(
          a, nbr, fl, slot, pbuf[src], kbuf[src], last ? a.p : pbuf[src ^ 1], kbuf[src ^ 1],
          op.surface_pressure_donors_all, last);
    }
    TSPH_CHECK_LAUNCH();
  }
  if (nw > 0) k_step_wall<D, S, NbSlots><<<grid_for(nw), kBlock, 0, stream>>>(a, nbr, fl, slot);
  if (op.boundary_consistent_density && nw > 0 && n > 0) {
    k_step_wall_mass<D, S, NbSlots><<<grid_for(n), kBlock, 0, stream>>>(a, nbr, fl, slot);
    k_step_wall<D, S, NbSlots><<<grid_for(nw), kBlock, 0, stream>>>(a, nbr, fl, slot);
  }
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// Density (+ EOS) and wall sweeps over the slot lists: the first two sweeps of
// a step on the cell-tile path, whose force sweep runs on the tiles.
template <int D, typename S>
int queue_slot_density_wall(tsph_sim* sim, int64_t k, hipStream_t stream) {
  const int slot = (int)(k & 1);
  SimFlags* fl = sim->flags.as<SimFlags>();
  WcsphArgs<S> a = sim_args<D, S>(sim);
  const NbSlots nbr = sim_slots(sim);
  const int r = (int)(k % kRing);
  launch_density<D, S, false>(a.shepard, grid_xcd(sim->n), 0, stream, a, nbr, fl, slot, 1, HaloView{}, NbHalo{});
  TSPH_CHECK_LAUNCH();
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][1], stream));
  TSPH_TRY((queue_wall_stage<D, S>(sim, a, nbr, fl, slot, stream)));
  return TSPH_OK;
}
template int queue_slot_density_wall<2, float>(tsph_sim*, int64_t, hipStream_t);
template int queue_slot_density_wall<2, double>(tsph_sim*, int64_t, hipStream_t);

// The slot-list sweeps of one step (used where no cell-tile path exists).
template <int D, typename S>
int queue_list_sweeps(tsph_sim* sim, int64_t k, bool meters, hipStream_t stream) {
  tsph_ctx* ctx = sim->ctx;
  const int64_t n = sim->n, nw = sim->nw;
  const int slot = (int)(k & 1);
  SimFlags* fl = sim->flags.as<SimFlags>();
  WcsphArgs<S> a = sim_args<D, S>(sim);
  const NbSlots nbr = sim_slots(sim);
  const unsigned gx = grid_xcd(n);
  const int r = (int)(k % kRing);
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][0], stream));
  // halo mode: the density and force sweeps stage their tiles' halos
  constexpr bool kHalo2 = true;
  const bool halo = sim->tiles == 2 && sim->halo_ok;
  if constexpr (D == 3) {
    if (halo) {
      TSPH_TRY((allow_halo_lds<k_step_density<D, S, NbSlots, true, false>>()));
      TSPH_TRY((allow_halo_lds<k_step_force<D, S, true, NbSlots, true>>()));
      TSPH_TRY((allow_halo_lds<k_step_force<D, S, false, NbSlots, true>>()));
    }
  }
  const HaloView hv = halo ? sim_halo(sim) : HaloView{};
  const NbHalo hl = halo ? sim_halo_lists(sim) : NbHalo{};
  const size_t halo_density_bytes = halo ? (size_t)sim->halo_max * (D + 1) * sizeof(double) : 0;
  const size_t halo_force_bytes = halo ? (size_t)sim->halo_max * sizeof(HaloRec<D, S>) : 0;
  if (sim->prm.op.diffusion_scale > 0.0) {
    if (sizeof(S) == 4) a.rho_lo = sim->rho_lo.as<S>();
    WcsphArgs<S> a0 = a;
    a0.rho = sim->rho_raw.as<S>();
    if (halo)
      launch_density<D, S, kHalo2>(a0.shepard, gx, halo_density_bytes, stream, a0, nbr, fl, slot, 0, hv, hl);
    else
      launch_density<D, S, false>(a0.shepard, gx, 0, stream, a0, nbr, fl, slot, 0, hv, hl);
    if (sim->prm.op.diffusion_form == 1) {
      k_step_gradient<D, S, NbSlots><<<gx, kBlock, 0, stream>>>(a0, nbr, sim->grad.as<S>(), fl, slot);
      k_step_diffusion<D, S, 1, NbSlots><<<gx, kBlock, 0, stream>>>(a, nbr, a0.rho, sim->grad.as<S>(), fl, slot);
    } else {
      k_step_diffusion<D, S, 0, NbSlots><<<gx, kBlock, 0, stream>>>(a, nbr, a0.rho, nullptr, fl, slot);
    }
  } else if (halo) {
    launch_density<D, S, kHalo2>(a.shepard, gx, halo_density_bytes, stream, a, nbr, fl, slot, 1, hv, hl);
  } else {
    launch_density<D, S, false>(a.shepard, gx, 0, stream, a, nbr, fl, slot, 1, hv, hl);
  }
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][1], stream));
  TSPH_TRY((queue_wall_stage<D, S>(sim, a, nbr, fl, slot, stream)));
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
  st.block_max = sim->block_max.as<unsigned long long>();
  st.gravity = sim->prm.op.gravity;
  if (D == 3 && sim->prm.obstacle_clamp) {
    st.obs_clamp = 1;
    for (int d = 0; d < 3; ++d) { st.obs_lo[d] = sim->prm.obstacle_lo[d]; st.obs_hi[d] = sim->prm.obstacle_hi[d]; }
    st.obs_eps = sim->prm.obstacle_eps;
    st.clamp_imp = sim->clamp_acc.as<double>();
    st.clamp_cnt = (unsigned long long*)(sim->clamp_acc.as<double>() + 3);
  }
  if (sim->obstacle && nw > 0) {
    // loads read the state before this step's integration (the same pr2 / vol
    // the force sweep uses); the force sweep writes the other pos / vel side
    const int64_t lo = n + sim->prm.obstacle_begin, hi = n + sim->prm.obstacle_end;
    const double* c = sim->prm.obstacle_centre;
    k_step_loads<D, S, NbSlots><<<grid_for(nw), kBlock, 0, stream>>>(a, nbr,
                                                             sim->orig.as<int64_t>(), lo, hi, c[0], c[1], c[2], fl,
                                                             slot, sim->load_part.as<double>());
    k_reduce_loads<<<1, 256, 0, stream>>>(sim->load_part.as<double>(), grid_for(nw), sim->loads_out.as<double>(),
                                          fl, slot);
    TSPH_CHECK_LAUNCH();
  }
  if (halo && meters)
    k_step_force<D, S, true, NbSlots, kHalo2><<<gx, kBlock, halo_force_bytes, stream>>>(
        a, nbr, sim->acc.as<S>(), st, fl, slot, 0.5 * sim->prm.skin, hv, hl);
  else if (halo)
    k_step_force<D, S, false, NbSlots, kHalo2><<<gx, kBlock, halo_force_bytes, stream>>>(
        a, nbr, sim->acc.as<S>(), st, fl, slot, 0.5 * sim->prm.skin, hv, hl);
  else if (meters)
    k_step_force<D, S, true, NbSlots><<<gx, kBlock, 0, stream>>>(a, nbr, sim->acc.as<S>(), st, fl, slot,
                                                                0.5 * sim->prm.skin, hv, hl);
  else
    k_step_force<D, S, false, NbSlots><<<gx, kBlock, 0, stream>>>(a, nbr, sim->acc.as<S>(), st, fl, slot,
                                                                 0.5 * sim->prm.skin, hv, hl);
  if (sim->tstep) TSPH_CHECK_HIP(hipEventRecord(sim->tev[r][3], stream));
  if (meters) k_reduce_sums<<<1, 256, 0, stream>>>(sim->block_sums.as<double>(), gx, sim->sums_out.as<double>(), fl, slot);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

static DtCtl sim_dtctl(const tsph_sim* sim) {
  const tsph_sim_params& p = sim->prm;
  DtCtl c{};
  c.adaptive = p.adaptive;
  c.dt_cfg = p.op.dt;
  c.target_time = p.target_time;
  c.tol = p.adaptive_tol > 0.0 ? p.adaptive_tol : 1.0e-12 * (p.target_time > 1.0 ? p.target_time : 1.0);
  c.free_last = p.adaptive_free_last;
  c.t0 = p.initial_time;
  c.dt0 = p.adaptive && p.initial_dt > 0.0 ? p.initial_dt : p.op.dt;
  c.courant_num = p.courant_num;
  c.c0 = p.op.sound_speed;
  c.dt_limit = p.dt_limit;
  c.dt_safety = p.dt_safety;
  c.dt_growth = p.dt_growth;
  return c;
}

// the time state before the first step (a resumed run starts from the
// elapsed time and step-size variable it was saved with)
__global__ void k_init_time(SimFlags* fl, DtCtl c) {
  fl->dt = c.dt0;
  fl->last_dt = c.dt0;
  fl->elapsed = c.t0;
  fl->final = c.adaptive && c.t0 >= c.target_time - c.tol ? 1ull : 0ull;
  const double left = c.target_time - c.t0;
  fl->step_dt = fl->final ? 0.0 : (c.adaptive && !c.free_last && left < c.dt0 ? left : c.dt0);
}

// Queue one step (density [+diffusion], wall, force+integrate) reading the
// stale slot of step `k`.
template <int D, typename S>
int sim_queue_step(tsph_sim* sim, int64_t k, bool meters, hipStream_t stream) {
  tsph_ctx* ctx = sim->ctx;
  const int64_t n = sim->n, nw = sim->nw;
  const int slot = (int)(k & 1);
  SimFlags* fl = sim->flags.as<SimFlags>();
  WcsphArgs<S> a = sim_args<D, S>(sim);
  const unsigned gx = grid_xcd(n);
  const int r = (int)(k % kRing);
  // sweep timing on every timing-th step (events cost a little on each step they time)
  sim->tstep = sim->timing > 0 && k % sim->timing == 0;
  sim->timed[r] = sim->tstep;
  if (meters) k_reset_meters<<<1, 64, 0, stream>>>(fl, slot);
  bool tiled = false;
  if constexpr (D == 2) {
    if (sim->tiles == 1) {
      TSPH_TRY((sim_queue_step_tiles<D, S>(sim, k, meters, stream)));
      const int64_t nblocks = sim->ntiles_f;   // partials are indexed by tile
      if (meters) k_reduce_sums<<<1, 256, 0, stream>>>(sim->block_sums.as<double>(), nblocks, sim->sums_out.as<double>(), fl, slot);
      TSPH_CHECK_LAUNCH();
      tiled = true;
    }
  }
  if (!tiled) TSPH_TRY((queue_list_sweeps<D, S>(sim, k, meters, stream)));
  sim->cur ^= 1;
  // record this step for the host: packed on the device, copied to the
  // pinned ring on a side stream so that the copy overlaps the next step
  StepRecord* rec_dev = sim->rec_dev.as<StepRecord>() + r;
  k_pack_record<<<1, kPackThreads, 0, stream>>>(fl, sim->sums_out.as<double>(), slot, meters ? 1 : 0, rec_dev,
                                                load_rows(sim) ? sim->loads_out.as<double>() : nullptr,
                                                load_rows(sim) ? sim->clamp_acc.as<double>() : nullptr,
                                                tiled ? nullptr : sim->block_max.as<unsigned long long>(),
                                                tiled ? 0 : (int64_t)gx, sim_dtctl(sim));
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipEventRecord(sim->pev[r], stream));
  TSPH_CHECK_HIP(hipStreamWaitEvent(sim->rec_stream, sim->pev[r], 0));
  TSPH_CHECK_HIP(hipMemcpyAsync(&sim->ring[r], rec_dev, sizeof(StepRecord), hipMemcpyDeviceToHost, sim->rec_stream));
  TSPH_CHECK_HIP(hipEventRecord(sim->ev[r], sim->rec_stream));
  sim->ring_step[r] = k;
  return TSPH_OK;
}

// Consume the record of step k: returns 1 if the step after it must be redone
// (its list went stale), 0 otherwise.
int sim_consume(tsph_sim* sim, int64_t k, bool meters, bool* stale_after, bool* skipped) {
  const int r = (int)(k % kRing);
  TSPH_CHECK_HIP(hipEventSynchronize(sim->ev[r]));
  StepRecord* rec = &sim->ring[r];
  *stale_after = rec->stale_out != 0;
  if (skipped) *skipped = rec->stale_in != 0;
  if (rec->stale_in) return TSPH_OK;   // skipped: nothing to record
  if (sim->finished) return TSPH_OK;   // after an adaptive run's final sample: no-op steps
  sim->steps_done = k + 1;
  if (sim->timed[r]) {
    for (int q = 0; q < 3; ++q) {
      float ms = 0.0f;
      TSPH_CHECK_HIP(hipEventElapsedTime(&ms, sim->tev[r][q], sim->tev[r][q + 1]));
      sim->sweep_ms[q] += ms;
    }
    sim->sweep_count++;
  }
  sim->escaped_total = (int64_t)rec->escaped;
  const double rho_max = ord_dec_host(rec->maxes[MM_PEAK_RHO]);
  const double spd = ord_dec_host(rec->maxes[MM_PEAK_SPEED]);
  sim->peak_rho = rho_max > sim->peak_rho ? rho_max : sim->peak_rho;
  sim->last_speed = spd;
  sim->last_step_dt = rec->step_dt;
  sim->peak_speed = spd > sim->peak_speed ? spd : sim->peak_speed;
  if (load_rows(sim)) {
    sim->load_rows.push_back((double)k);
    for (int i = 0; i < 6; ++i) sim->load_rows.push_back(rec->loads[i]);
    for (int i = 0; i < 3; ++i) sim->load_rows.push_back(rec->clamp_imp[i]);
    sim->load_rows.push_back((double)rec->clamp_count);
    sim->load_rows.push_back(rec->step_dt);
  }
  if (meters) {
    sim->rows.push_back((double)k);
    for (int i = 0; i < MS_NSUM; ++i) sim->rows.push_back(rec->sums[i]);
    for (int i = 0; i < MM_NMAX; ++i) sim->rows.push_back(ord_dec_host(rec->maxes[i]));
    for (int i = 0; i < MN_NMIN; ++i) sim->rows.push_back(ord_dec_host(rec->mins[i]));
    sim->rows.push_back(rec->elapsed);
    sim->rows.push_back(rec->last_dt);
    sim->rows.push_back(rec->step_dt);
    sim->rows.push_back(rec->dt);
  }
  if (rec->final) sim->finished = true;
  return TSPH_OK;
}

template <int D, typename S>
int sim_run(tsph_sim* sim, int64_t nsteps, int64_t meter_every, hipStream_t stream) {
  const int64_t start = sim->steps_done;
  const int64_t end = start + nsteps;
  auto rebuild = [&]() -> int {
    sim->rebuild_forced = false;       // a counted build re-anchors every row
    if constexpr (D == 2) {
      if (sim->tiles == 1) return sim_rebuild_tiles<D, S>(sim, stream);
    }
    return sim_rebuild<D, S>(sim, stream, true);
  };
  if (sim->finished) return TSPH_OK;
  if (!sim->built || sim->rebuild_pending) {
    TSPH_TRY(rebuild());
    sim->rebuild_pending = false;
  }
  // steps in flight (queued but not yet consumed)
  int64_t next = start;        // next step index to queue
  int64_t done = start;        // next step index to consume
  const int depth = 2;
  auto is_meter = [&](int64_t k) { return meter_every > 0 && (k % meter_every) == 0; };
  while (done < end) {
    while (next < end && next - done < depth) {
      TSPH_TRY((sim_queue_step<D, S>(sim, next, is_meter(next), stream)));
      ++next;
    }
    bool stale_after = false, skipped = false;
    TSPH_TRY(sim_consume(sim, done, is_meter(done), &stale_after, &skipped));
    // a step skipped on the device (its build's lists did not fit) is run
    // again after the next build
    if (!skipped) ++done;
    if (sim->finished) {
      // the final sample is done; steps queued after it integrate with dt = 0
      for (int64_t k = done; k < next; ++k) TSPH_CHECK_HIP(hipEventSynchronize(sim->ev[k % kRing]));
      break;
    }
    if (stale_after || skipped) {
      // steps done..next-1 were skipped on the device; drain them, rebuild, redo
      for (int64_t k = skipped ? done + 1 : done; k < next; ++k) {
        bool dummy;
        TSPH_TRY(sim_consume(sim, k, false, &dummy));
      }
      if (done >= end) {
        // nothing left to run: leave the last step's input and output sides
        // as they are (the gauges read its input side) and re-bin first
        // thing in the next run
        sim->rebuild_pending = true;
        break;
      }
      TSPH_TRY(rebuild());
      next = done;
    }
  }
  return TSPH_OK;
}

}  // namespace tsph

extern "C" {

int tsph_sim_create(const tsph_sim_params* prm, const double* pos, const double* vel, const void* mass,
                    hipStream_t stream, tsph_sim** out) {
  TSPH_REQUIRE(prm && out && pos && vel && mass, TSPH_ERR_ARG, "NULL argument");
  TSPH_REQUIRE(prm->op.dim == 2 || prm->op.dim == 3, TSPH_ERR_ARG, "dim must be 2 or 3");
  TSPH_REQUIRE(prm->op.precision == TSPH_F64 || prm->op.precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(prm->op.n_fluid > 0 && prm->op.n_wall >= 0, TSPH_ERR_ARG, "need at least one fluid particle");
  TSPH_REQUIRE(prm->skin > 0.0, TSPH_ERR_ARG, "the resident stepper needs a positive Verlet skin");
  TSPH_REQUIRE(prm->op.h > 0.0 && prm->op.dt > 0.0, TSPH_ERR_ARG, "h and dt must be positive");
  TSPH_REQUIRE(prm->op.reaction_style == 0 && prm->op.has_wall_accel == 0, TSPH_ERR_ARG,
               "the resident stepper runs static walls (obstacle loads come through obstacle_begin/end)");
  TSPH_REQUIRE(prm->obstacle_begin >= 0 && prm->obstacle_begin <= prm->obstacle_end &&
               prm->obstacle_end <= prm->op.n_wall, TSPH_ERR_ARG, "obstacle rows outside the wall rows");
  TSPH_REQUIRE(!prm->obstacle_clamp || prm->op.dim == 3, TSPH_ERR_ARG, "the obstacle clamp is 3-D");
  TSPH_REQUIRE(prm->op.wrap_width == 0.0, TSPH_ERR_ARG, "the resident stepper has no x-periodic strip");
  TSPH_REQUIRE(prm->op.dim == 2 || !(prm->op.density_renormalization || prm->op.boundary_consistent_density ||
                                     prm->op.surface_pressure_closure),
               TSPH_ERR_ARG, "density renormalisation, boundary-consistent density and the pressure closure "
               "are 2-D rhs branches");
  TSPH_REQUIRE(!prm->op.density_renormalization || prm->op.lattice_kernel_sum > 0.0, TSPH_ERR_ARG,
               "density_renormalization needs a positive lattice_kernel_sum");
  if (prm->adaptive)
    TSPH_REQUIRE(prm->target_time >= 0.0 && prm->courant_num > 0.0 && prm->dt_limit > 0.0 &&
                 prm->dt_safety > 0.0 && prm->dt_growth > 0.0, TSPH_ERR_ARG,
                 "adaptive step: target_time >= 0 and positive courant_num, dt_limit, dt_safety, dt_growth");
  if (prm->obstacle_clamp)
    for (int d = 0; d < 3; ++d)
      TSPH_REQUIRE(prm->obstacle_lo[d] < prm->obstacle_hi[d] && prm->obstacle_eps >= 0.0, TSPH_ERR_ARG,
                   "obstacle box must have lo < hi and eps >= 0");
  tsph_sim* sim = new tsph_sim();
  sim->prm = *prm;
  sim->dim = prm->op.dim;
  sim->prec = prm->op.precision;
  sim->n = prm->op.n_fluid; sim->nw = prm->op.n_wall; sim->nt = sim->n + sim->nw;
  sim->elt = sim->prec == TSPH_F64 ? 8 : 4;
  int st = tsph_ctx_create(&sim->ctx);
  if (st != TSPH_OK) { delete sim; return st; }
  const int D = sim->dim;
  const int64_t nt = sim->nt, n = sim->n;
  const size_t e = sim->elt;
  int rc = TSPH_OK;
#define ENS(buf, bytes) if (rc == TSPH_OK) rc = (buf).ensure(bytes)
  ENS(sim->pos[0], 8 * nt * D); ENS(sim->pos[1], 8 * nt * D);
  ENS(sim->vel[0], 8 * nt * D); ENS(sim->vel[1], 8 * nt * D);
  // FP32 runs keep each mass as two floats: hi in [0, nt), lo in [nt, 2 nt)
  ENS(sim->mass, e * nt * (e == 4 ? 2 : 1)); ENS(sim->mass2, e * nt * (e == 4 ? 2 : 1));
  ENS(sim->orig, 8 * nt); ENS(sim->orig2, 8 * nt);
  ENS(sim->rho, e * nt); ENS(sim->p, e * nt); ENS(sim->pr2, e * nt); ENS(sim->vol, e * nt);
  if (e == 4) { ENS(sim->aux, 4 * e * nt); ENS(sim->p_lo, e * nt); ENS(sim->gsum, e * n * D + 4); }
  if (prm->op.surface_pressure_closure) {
    const size_t sb = (e * (size_t)(prm->op.n_fluid + 1) + 255) / 256 * 256;
    const size_t kb = (sizeof(int) * (size_t)(prm->op.n_fluid + 1) + 255) / 256 * 256;
    ENS(sim->closure, 2 * sb + 2 * kb);
  }
  ENS(sim->rho_raw, e * nt); ENS(sim->rho_lo, e * nt); ENS(sim->grad, e * n * D); ENS(sim->acc, e * n * D);
  ENS(sim->anchor, 8 * n * D);
  ENS(sim->flags, sizeof(SimFlags));
  ENS(sim->block_sums, 8 * MS_NSUM * (int64_t)grid_xcd(n));
  ENS(sim->block_max, 8 * 4 * (int64_t)grid_xcd(n));
  ENS(sim->rec_dev, sizeof(StepRecord) * kRing);
  sim->obstacle = prm->obstacle_end > prm->obstacle_begin ? 1 : 0;
  ENS(sim->load_part, 8 * 6 * ((int64_t)grid_for(sim->nw) + 1));
  ENS(sim->loads_out, 8 * 8);
  ENS(sim->clamp_acc, 8 * 4);
  ENS(sim->sums_out, 8 * MS_NSUM);
  ENS(sim->new2old, 8 * nt);
  ENS(sim->cells_new, 8 * nt * 3);
#undef ENS
  if (rc != TSPH_OK) { tsph_sim_destroy(sim); return rc; }
  hipError_t he = hipHostMalloc((void**)&sim->ring, sizeof(StepRecord) * kRing, hipHostMallocDefault);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&sim->rec_stream, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipHostMalloc((void**)&sim->bin_outside, 64, hipHostMallocDefault);
  if (he != hipSuccess) { set_error("hipHostMalloc: %s", hipGetErrorString(he)); tsph_sim_destroy(sim); return TSPH_ERR_HIP; }
  memset(sim->bin_outside, 0, 64);
  sim->build_stats = reinterpret_cast<BuildStats*>(sim->bin_outside + 4);
  memset(sim->ring, 0, sizeof(StepRecord) * kRing);
  for (int r = 0; r < kRing; ++r) {
    he = hipEventCreateWithFlags(&sim->ev[r], hipEventDisableTiming);
    // pev only orders the record copy (a kernel on this device) after the
    // packing kernel: a device-scope release is enough, and a system-scope
    // one would write back the L2 after every step
    if (he == hipSuccess)
      he = hipEventCreateWithFlags(&sim->pev[r], hipEventDisableTiming | hipEventDisableSystemFence);
    if (he != hipSuccess) { set_error("hipEventCreate: %s", hipGetErrorString(he)); tsph_sim_destroy(sim); return TSPH_ERR_HIP; }
  }
  for (int r = 0; r < kRing; ++r)
    for (int q = 0; q < 4; ++q) {
      // timing only: no system-scope fence, whose L2 writeback would be
      // counted in the sweep it brackets
      he = hipEventCreateWithFlags(&sim->tev[r][q], hipEventDisableSystemFence);
      if (he != hipSuccess) { set_error("hipEventCreate: %s", hipGetErrorString(he)); tsph_sim_destroy(sim); return TSPH_ERR_HIP; }
    }
  // initial state: caller order; velocities of wall rows are zero
  he = hipMemcpyAsync(sim->pos[0].ptr, pos, 8 * nt * D, hipMemcpyDeviceToDevice, stream);
  if (he == hipSuccess) he = hipMemsetAsync(sim->vel[0].ptr, 0, 8 * nt * D, stream);
  if (he == hipSuccess) he = hipMemcpyAsync(sim->vel[0].ptr, vel, 8 * n * D, hipMemcpyDeviceToDevice, stream);
  if (he == hipSuccess) he = hipMemcpyAsync(sim->mass.ptr, mass, e * nt, hipMemcpyDeviceToDevice, stream);
  if (he == hipSuccess && e == 4) he = hipMemsetAsync(sim->mass.as<float>() + nt, 0, e * nt, stream);
  if (he == hipSuccess) he = hipMemsetAsync(sim->flags.ptr, 0, sizeof(SimFlags), stream);
  if (he == hipSuccess) he = hipMemsetAsync(sim->loads_out.ptr, 0, 8 * 8, stream);
  if (he == hipSuccess) he = hipMemsetAsync(sim->clamp_acc.ptr, 0, 8 * 4, stream);
  if (he != hipSuccess) { set_error("sim init: %s", hipGetErrorString(he)); tsph_sim_destroy(sim); return TSPH_ERR_HIP; }
  k_iota<<<grid_for(nt), kBlock, 0, stream>>>(sim->orig.as<int64_t>(), nt);
  k_init_time<<<1, 1, 0, stream>>>(sim->flags.as<SimFlags>(), sim_dtctl(sim));
  // runs default to the slot lists with LDS halos (measured fastest on
  // MI355X: 2-D dam break at 1.6 M fluid rows, 3-D box case at 20 M, FP32
  // and FP64); the cell tiles (2-D) and the plain lists stay selectable
  sim->tile_dim[0] = 8; sim->tile_dim[1] = 4;
  sim->tiles = 2;
  *out = sim;
  return TSPH_OK;
}

int tsph_sim_step(tsph_sim* sim, int64_t nsteps, int64_t meter_every, hipStream_t stream) {
  TSPH_REQUIRE(sim && nsteps >= 0, TSPH_ERR_ARG, "bad arguments");
  if (nsteps == 0) return TSPH_OK;
  if (sim->dim == 2)
    return sim->prec == TSPH_F64 ? sim_run<2, double>(sim, nsteps, meter_every, stream)
                                 : sim_run<2, float>(sim, nsteps, meter_every, stream);
  return sim->prec == TSPH_F64 ? sim_run<3, double>(sim, nsteps, meter_every, stream)
                               : sim_run<3, float>(sim, nsteps, meter_every, stream);
}

int tsph_sim_set_masses_f64(tsph_sim* sim, const double* mass, hipStream_t stream) {
  TSPH_REQUIRE(sim && mass, TSPH_ERR_ARG, "sim and mass must not be NULL");
  const int64_t nt = sim->nt;
  if (sim->prec == TSPH_F64)
    k_set_masses<double><<<grid_for(nt), kBlock, 0, stream>>>(mass, sim->orig.as<int64_t>(), nt,
                                                              sim->mass.as<double>(), 0);
  else
    k_set_masses<float><<<grid_for(nt), kBlock, 0, stream>>>(mass, sim->orig.as<int64_t>(), nt,
                                                             sim->mass.as<float>(), mass_lo_offset(sim));
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

int tsph_sim_read(tsph_sim* sim, double* pos, double* vel, void* rho, void* p, void* acc, hipStream_t stream) {
  TSPH_REQUIRE(sim, TSPH_ERR_ARG, "sim must not be NULL");
  const int64_t n = sim->n;
  const unsigned g = grid_for(n);
  const double* P = sim->pos[sim->cur].as<double>();
  const double* V = sim->vel[sim->cur].as<double>();
  if (sim->dim == 2) {
    if (sim->prec == TSPH_F64)
      k_unpermute<2, double><<<g, kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, P, V, sim->rho.as<double>(), sim->p.as<double>(), sim->acc.as<double>(), pos, vel, (double*)rho, (double*)p, (double*)acc);
    else
      k_unpermute<2, float><<<g, kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, P, V, sim->rho.as<float>(), sim->p.as<float>(), sim->acc.as<float>(), pos, vel, (float*)rho, (float*)p, (float*)acc);
  } else {
    if (sim->prec == TSPH_F64)
      k_unpermute<3, double><<<g, kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, P, V, sim->rho.as<double>(), sim->p.as<double>(), sim->acc.as<double>(), pos, vel, (double*)rho, (double*)p, (double*)acc);
    else
      k_unpermute<3, float><<<g, kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, P, V, sim->rho.as<float>(), sim->p.as<float>(), sim->acc.as<float>(), pos, vel, (float*)rho, (float*)p, (float*)acc);
  }
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

int tsph_sim_time(tsph_sim* sim, double* out, hipStream_t stream) {
  TSPH_REQUIRE(sim && out, TSPH_ERR_ARG, "bad arguments");
  const SimFlags* fl = sim->flags.as<SimFlags>();
  double t[4];
  TSPH_CHECK_HIP(hipMemcpyAsync(&t[0], &fl->elapsed, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(&t[1], &fl->dt, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(&t[2], &fl->step_dt, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(&t[3], &fl->last_dt, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  for (int k = 0; k < 4; ++k) out[k] = t[k];
  return TSPH_OK;
}

int tsph_sim_stats(tsph_sim* sim, int64_t* counters, double* peaks) {
  TSPH_REQUIRE(sim, TSPH_ERR_ARG, "sim must not be NULL");
  if (counters) {
    counters[0] = sim->steps_done;
    counters[1] = sim->rebuilds;
    counters[2] = sim->escaped_total + sim->escaped_extra;
    counters[3] = sim->ncells_dense;
    counters[4] = (int64_t)(sim->rows.size() / kMeterWidth);
    counters[5] = sim->finished ? 1 : 0;
  }
  if (peaks) { peaks[0] = sim->peak_rho; peaks[1] = sim->peak_speed; }
  return TSPH_OK;
}

int tsph_sim_slot_width(tsph_sim* sim, int64_t width, int64_t* current) {
  TSPH_REQUIRE(sim, TSPH_ERR_ARG, "sim must not be NULL");
  if (current) *current = sim->slot_K;
  if (width >= 0) sim->slot_K_next = (width + 3) / 4 * 4;   // the lists in place keep their width
  return TSPH_OK;
}

int tsph_sim_layout(tsph_sim* sim, int64_t* out) {
  TSPH_REQUIRE(sim && out, TSPH_ERR_ARG, "bad arguments");
  out[0] = sim->tiles;
  out[1] = sim->halo_ok ? (int64_t)sim->build_stats->used : 0;
  out[2] = sim->halo_ok ? sim->halo_tiles : 0;
  out[3] = sim->halo_ok ? (int64_t)sim->build_stats->hmax : 0;
  return TSPH_OK;
}

int tsph_sim_meters(tsph_sim* sim, double* rows, int64_t max_rows, int64_t* nrows) {
  TSPH_REQUIRE(sim && nrows, TSPH_ERR_ARG, "bad arguments");
  const int64_t width = kMeterWidth;
  const int64_t have = (int64_t)sim->rows.size() / width;
  *nrows = have;
  if (rows) {
    const int64_t m = have < max_rows ? have : max_rows;
    memcpy(rows, sim->rows.data(), sizeof(double) * width * m);
  }
  return TSPH_OK;
}

int tsph_sim_loads(tsph_sim* sim, double* rows, int64_t max_rows, int64_t* nrows) {
  TSPH_REQUIRE(sim && nrows, TSPH_ERR_ARG, "bad arguments");
  const int64_t width = 12;
  const int64_t have = (int64_t)sim->load_rows.size() / width;
  *nrows = have;
  if (rows) {
    const int64_t m = have < max_rows ? have : max_rows;
    memcpy(rows, sim->load_rows.data(), sizeof(double) * width * m);
  }
  return TSPH_OK;
}

int tsph_sim_configure(tsph_sim* sim, int tiles, int tile_x, int tile_y) {
  TSPH_REQUIRE(sim, TSPH_ERR_ARG, "sim must not be NULL");
  TSPH_REQUIRE(tiles >= 0 && tiles <= 2, TSPH_ERR_ARG, "layout must be 0 (lists), 1 (cell tiles) or 2 (halos)");
  TSPH_REQUIRE(tiles != 1 || sim->dim == 2, TSPH_ERR_ARG, "cell-tile sweeps are 2-D");
  TSPH_REQUIRE(tiles != 1 || !sim->obstacle, TSPH_ERR_ARG, "obstacle loads run on the slot-list sweeps");
  TSPH_REQUIRE(tiles != 1 || !(sim->prm.op.diffusion_scale > 0.0), TSPH_ERR_ARG,
               "density diffusion runs on the slot-list sweeps");
  TSPH_REQUIRE(tiles != 1 || (tile_x >= 1 && tile_y >= 1 && (tile_x + 2) * (tile_y + 2) <= 18 * 6 &&
                              tile_x * tile_y <= 64),
               TSPH_ERR_ARG, "tile %d x %d out of range ((tx+2)(ty+2) <= 108, tx*ty <= 64)", tile_x, tile_y);
  sim->tiles = tiles;
  sim->halo_ok = false;
  sim->bgrid_ok = false;
  if (tiles == 1) { sim->tile_dim[0] = tile_x; sim->tile_dim[1] = tile_y; }
  sim->built = false;   // next step rebuilds with the new layout
  return TSPH_OK;
}

int tsph_sim_timing(tsph_sim* sim, int enable, double* sweep_ms, int64_t* steps) {
  TSPH_REQUIRE(sim, TSPH_ERR_ARG, "sim must not be NULL");
  if (sweep_ms) for (int q = 0; q < 3; ++q) sweep_ms[q] = sim->sweep_ms[q];
  if (steps) *steps = sim->sweep_count;
  if (enable >= 0) {
    sim->timing = enable;
    for (int q = 0; q < 3; ++q) sim->sweep_ms[q] = 0.0;
    sim->sweep_count = 0;
  }
  return TSPH_OK;
}

int tsph_sim_count_pairs(tsph_sim* sim, double radius, int64_t* pairs, hipStream_t stream) {
  TSPH_REQUIRE(sim && pairs && radius > 0.0, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(sim->built, TSPH_ERR_ARG, "no cell binning yet: step the simulation first");
  if (sim->rebuild_pending) TSPH_TRY(tsph_sim_step(sim, 0, 0, stream));   // re-bin the current state
  tsph_ctx* ctx = sim->ctx;
  TSPH_TRY(ctx->s_red.ensure(sizeof(unsigned long long)));
  unsigned long long* cnt = ctx->s_red.as<unsigned long long>();
  TSPH_CHECK_HIP(hipMemsetAsync(cnt, 0, 8, stream));
  const double r2 = radius * radius;
  TSPH_REQUIRE(radius <= 2.0 * sim->prm.op.h + sim->prm.skin, TSPH_ERR_ARG,
               "radius exceeds the neighbour-list build radius");
  if (sim->tiles == 1) {
    TSPH_TRY(tile_count_pairs(sim, radius, cnt, stream));
  } else {
    if (sim->dim == 2)
      k_count_close<2><<<grid_for(sim->n), kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), sim->n,
                                                                sim_cells(sim), r2, cnt);
    else
      k_count_close<3><<<grid_for(sim->n), kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), sim->n,
                                                                sim_cells(sim), r2, cnt);
    TSPH_CHECK_LAUNCH();
  }
  unsigned long long h = 0;
  TSPH_CHECK_HIP(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  *pairs = (int64_t)h;
  return TSPH_OK;
}

int tsph_sim_destroy(tsph_sim* sim) {
  if (!sim) return TSPH_OK;
  if (sim->rec_stream) {                  // record copies still in flight read rec_dev
    (void)hipStreamSynchronize(sim->rec_stream);
    (void)hipStreamDestroy(sim->rec_stream);
  }
  DevBuf* bufs[] = {&sim->pos[0], &sim->pos[1], &sim->vel[0], &sim->vel[1], &sim->mass, &sim->mass2,
                    &sim->orig, &sim->orig2, &sim->rho, &sim->p, &sim->pr2, &sim->vol, &sim->aux, &sim->p_lo, &sim->gsum, &sim->rho_raw, &sim->rho_lo,
                    &sim->grad, &sim->acc, &sim->anchor, &sim->flags, &sim->block_sums, &sim->block_max, &sim->rec_dev, &sim->sums_out,
                    &sim->new2old, &sim->cells_new, &sim->bin_cnt, &sim->bin_start,
                    &sim->bin_part, &sim->bin_keys, &sim->bin_tmp, &sim->bin_cells, &sim->bin_flag,
                    &sim->bin_bounds, &sim->cell_R, &sim->slot_idx,
                    &sim->slot_cnt, &sim->slot_max, &sim->slot_posf, &sim->ranges,
                    &sim->tiles_f, &sim->tiles_w, &sim->tkeys, &sim->tkeys_sorted, &sim->tidx, &sim->tflags,
                    &sim->load_part, &sim->loads_out, &sim->clamp_acc, &sim->closure, &sim->shift_buf,
                    &sim->shift_tmp, &sim->halo_rows, &sim->halo_cnt, &sim->halo_idx, &sim->halo_hmax};
  for (DevBuf* b : bufs) b->release();
  for (int r = 0; r < kRing; ++r) {
    if (sim->ev[r]) (void)hipEventDestroy(sim->ev[r]);
    if (sim->pev[r]) (void)hipEventDestroy(sim->pev[r]);
  }
  for (int r = 0; r < kRing; ++r)
    for (int q = 0; q < 4; ++q) if (sim->tev[r][q]) (void)hipEventDestroy(sim->tev[r][q]);
  if (sim->ring) (void)hipHostFree(sim->ring);
  if (sim->bin_outside) (void)hipHostFree(sim->bin_outside);
  if (sim->ctx) tsph_ctx_destroy(sim->ctx);
  delete sim;
  return TSPH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- particle shifting
// particle_shifting.shift_positions (particle_shifting.py:88-224) on the
// stepper's fluid rows, over the fluid entries of the slot lists within 2h:
// C_i = V_i W(0) + sum_j V_j W_ij, grad C_i = sum_j V_j (C_j - C_i) grad_i W_ij,
// shift = -A grad C (A = strength h v_ref dt), tangential at the free surface
// (C_i < free_surface C_ref), capped at limit h.
namespace tsph {

template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_shift_conc(WcsphArgs<S> a, NbSlots nbr, double* __restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> acc;
  nbr.template visit<D>(i, false, [&](int64_t j) {
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double df = xi[d] - a.pos[j * D + d];
      r2 += df * df;
    }
    if (r2 > a.skip2d) return;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    acc.add((double)a.vol[j] * w);
  });
  c[i] = (double)a.vol[i] * a.kd.w0d + hi(acc);
}

struct ShiftArgs {
  double amp, cap, fs_c;     // A, limit h, free_surface * C_ref (<= 0: no free surface)
};

template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_shift_vec(WcsphArgs<S> a, NbSlots nbr, const double* __restrict__ c,
                                                      ShiftArgs sa, double* __restrict__ shift,
                                                      unsigned long long* __restrict__ stats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  const double ci = c[i];
  Sum<double> gr[D];
  nbr.template visit<D>(i, false, [&](int64_t j) {
    double df[D], r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = xi[d] - a.pos[j * D + d]; r2 += df[d] * df[d]; }
    if (r2 > a.skip2d) return;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    const double s = (double)a.vol[j] * (c[j] - ci);
#pragma unroll
    for (int d = 0; d < D; ++d) gr[d].add(s * (df[d] * g));
  });
  double gv[D], sh[D], nn = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) { gv[d] = hi(gr[d]); nn += gv[d] * gv[d]; sh[d] = -sa.amp * gv[d]; }
  const double norm = sqrt(nn);
  if (sa.fs_c > 0.0 && ci < sa.fs_c) {
    double nh[D], dot = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      nh[d] = norm > 1.0e-30 ? -gv[d] / norm : 0.0;
      dot += sh[d] * nh[d];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) sh[d] = sh[d] - dot * nh[d];
  }
  double m2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) m2 += sh[d] * sh[d];
  double mag = sqrt(m2);
  if (mag > sa.cap) {
    const double sc = sa.cap / (mag > 1.0e-300 ? mag : 1.0e-300);
#pragma unroll
    for (int d = 0; d < D; ++d) sh[d] = sh[d] * sc;
    atomicAdd(&stats[0], 1ull);
    mag = sa.cap;
  }
#pragma unroll
  for (int d = 0; d < D; ++d) shift[i * D + d] = sh[d];
  atomicMax(&stats[1], ord_enc(mag));
}

// moves the rows and measures the displacement since the lists were built
template <int D>
__global__ void k_shift_move(double* __restrict__ pos, const double* __restrict__ shift,
                             const double* __restrict__ anchor, int64_t n, unsigned long long* __restrict__ stats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double x = pos[i * D + d] + shift[i * D + d];
    pos[i * D + d] = x;
    const double e = x - anchor[i * D + d];
    d2 += e * e;
  }
  atomicMax(&stats[2], ord_enc(sqrt(d2)));
}

// row values (w per row) to the caller's order and back, through the row ->
// caller map
template <typename S>
__global__ void k_rows_to_caller(const int64_t* __restrict__ orig, int64_t n, const S* __restrict__ src,
                                 S* __restrict__ dst, int w = 1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    for (int d = 0; d < w; ++d) dst[orig[i] * w + d] = src[i * w + d];
}
template <typename S>
__global__ void k_caller_to_rows(const int64_t* __restrict__ orig, int64_t n, const S* __restrict__ src,
                                 S* __restrict__ dst, int w = 1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    for (int d = 0; d < w; ++d) dst[i * w + d] = src[orig[i] * w + d];
}

// After a correction the host decides the Verlet test on the corrected
// positions.  When they are back within half the skin of the anchors, the
// stale flag the last step may have left on the device is cleared too, so the
// next step runs on the current lists, as the reference's pair cache would.
inline int clear_stale_unless_pending(tsph_sim* sim, hipStream_t stream) {
  if (sim->rebuild_pending) return TSPH_OK;
  SimFlags* fl = sim->flags.as<SimFlags>();
  TSPH_CHECK_HIP(hipMemsetAsync(&fl->stale[0], 0, sizeof(unsigned long long) * 2, stream));
  return TSPH_OK;
}

// A correction that needs the pairs within 2h of the current positions re-bins
// first when the last step left the lists stale, and carries the last
// evaluation's volumes, densities, pressures and accelerations into the new
// row order through the caller's order (tsph_sim_read reports them after
// the correction).
// That re-binning only completes the lists for the correction: it is not
// counted, and *rebuilt tells the caller to leave a rebuild pending, so the
// next step re-bins the corrected positions as the reference's pair cache
// does (same list builds, same anchors).
template <int D, typename S>
int rebuild_keep_volumes(tsph_sim* sim, hipStream_t stream, bool* rebuilt) {
  const int64_t n = sim->n;
  *rebuilt = false;
  if (sim->rebuild_pending) {
    // the lists went stale in the last step: re-bin, and carry the last
    // evaluation's volumes into the new row order
    TSPH_TRY(sim->shift_tmp.ensure(sizeof(S) * (size_t)(n * (3 + D) + 1)));
    S* carry = sim->shift_tmp.as<S>();
    S* const fields[3] = {sim->vol.as<S>(), sim->rho.as<S>(), sim->p.as<S>()};
    for (int f = 0; f < 3; ++f)
      k_rows_to_caller<S><<<grid_for(n), kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, fields[f], carry + f * n);
    k_rows_to_caller<S><<<grid_for(n), kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, sim->acc.as<S>(),
                                                            carry + 3 * n, D);
    TSPH_CHECK_LAUNCH();
    bool done = false;
    if constexpr (D == 2) {
      if (sim->tiles == 1) { TSPH_TRY((sim_rebuild_tiles<D, S>(sim, stream))); done = true; }
    }
    if (!done) TSPH_TRY((sim_rebuild<D, S>(sim, stream)));
    sim->rebuild_pending = false;
    sim->rebuilds--;
    sim->rebuild_forced = true;
    *rebuilt = true;
    for (int f = 0; f < 3; ++f)
      k_caller_to_rows<S><<<grid_for(n), kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, carry + f * n, fields[f]);
    k_caller_to_rows<S><<<grid_for(n), kBlock, 0, stream>>>(sim->orig.as<int64_t>(), n, carry + 3 * n,
                                                            sim->acc.as<S>(), D);
    TSPH_CHECK_LAUNCH();
  }
  return TSPH_OK;
}

// the percentile of c[0, n) with linear interpolation between order statistics
template <typename S>
int percentile_of(tsph_sim* sim, const double* c, double* sorted, int64_t n, double percentile, double* out,
                  hipStream_t stream) {
  size_t bytes = 0;
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, c, sorted, (int)n, 0, 64, stream));
  TSPH_TRY(sim->shift_tmp.ensure(bytes > sizeof(S) * (size_t)(n + 1) ? bytes : sizeof(S) * (size_t)(n + 1)));
  TSPH_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(sim->shift_tmp.ptr, bytes, c, sorted, (int)n, 0, 64, stream));
  const double q = (percentile < 0.0 ? 0.0 : (percentile > 100.0 ? 100.0 : percentile)) / 100.0;
  const double at = q * (double)(n - 1);
  const int64_t lo = (int64_t)floor(at), hi_i = lo + 1 < n ? lo + 1 : lo;
  double v[2];
  TSPH_CHECK_HIP(hipMemcpyAsync(&v[0], sorted + lo, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipMemcpyAsync(&v[1], sorted + hi_i, sizeof(double), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  *out = v[0] + (at - (double)lo) * (v[1] - v[0]);
  return TSPH_OK;
}

template <int D, typename S>
int sim_shift(tsph_sim* sim, const tsph_shift_params* prm, double* info, hipStream_t stream) {
  const int64_t n = sim->n;
  bool rebuilt = false;
  TSPH_TRY((rebuild_keep_volumes<D, S>(sim, stream, &rebuilt)));
  const size_t cb = (sizeof(double) * (size_t)(n + 1) + 255) / 256 * 256;
  TSPH_TRY(sim->shift_buf.ensure(2 * cb + (sizeof(double) * (size_t)n * D + 255) / 256 * 256 + 256));
  char* base = sim->shift_buf.as<char>();
  double* c = (double*)base;
  double* sorted = (double*)(base + cb);
  double* shift = (double*)(base + 2 * cb);
  unsigned long long* st = (unsigned long long*)(base + 2 * cb + (sizeof(double) * (size_t)n * D + 255) / 256 * 256);
  WcsphArgs<S> a = sim_args<D, S>(sim);
  const NbSlots nbr = sim_slots(sim);
  k_shift_conc<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, nbr, c);
  TSPH_CHECK_LAUNCH();
  // negative c_reference / v_ref: not given (the reference's None); zero is a value
  double c_ref = prm->c_reference;
  if (c_ref < 0.0) TSPH_TRY((percentile_of<S>(sim, c, sorted, n, prm->percentile, &c_ref, stream)));
  double v_ref = prm->v_ref;
  if (v_ref < 0.0) v_ref = sim->last_speed > prm->v_floor ? sim->last_speed : prm->v_floor;
  ShiftArgs sa;
  const double dt = prm->dt > 0.0 ? prm->dt : sim->last_step_dt;
  sa.amp = prm->strength * sim->prm.op.h * v_ref * dt;
  sa.cap = prm->limit * sim->prm.op.h;
  sa.fs_c = c_ref > 0.0 ? prm->free_surface * c_ref : 0.0;
  TSPH_CHECK_HIP(hipMemsetAsync(st, 0, 3 * sizeof(unsigned long long), stream));
  k_shift_vec<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, nbr, c, sa, shift, st);
  k_shift_move<D><<<grid_for(n), kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), shift,
                                                      sim->anchor.as<double>(), n, st);
  TSPH_CHECK_LAUNCH();
  unsigned long long h[3];
  TSPH_CHECK_HIP(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  const double disp = n > 0 ? ord_dec_host(h[2]) : 0.0;
  // the next step re-bins when the corrected positions are more than half the
  // skin from the anchors, the pair cache's own test
  sim->rebuild_pending = sim->rebuild_forced || disp > 0.5 * sim->prm.skin;
  TSPH_TRY(clear_stale_unless_pending(sim, stream));
  if (info) {
    info[0] = c_ref;
    info[1] = n > 0 ? ord_dec_host(h[1]) : 0.0;
    info[2] = n > 0 ? (double)h[0] / (double)n : 0.0;
    info[3] = v_ref;
  }
  return TSPH_OK;
}

// Transport-velocity correction (transport_velocity_correction.py:285-513),
// unit kernel correction: g_i = -2 sum_j V_j grad_i W_ij over fluid and wall
// neighbours within 2h (wall rows: V = wall_volume), shift = alpha h^2
// L(h^2 |g|^2) g, masked by the scope, capped at max_shift.
struct TvcArgs {
  double coef, slope, cap, wall_volume, h;
  int scope;                 // 0 all, 1 C >= c_thr, 2 |g| h <= g_tol
  double c_thr, g_tol;
};

template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_tvc_vec(WcsphArgs<S> a, NbSlots nbr, const double* __restrict__ c,
                                                    TvcArgs ta, double* __restrict__ shift,
                                                    unsigned long long* __restrict__ stats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> gr[D];
  nbr.template visit<D>(i, true, [&](int64_t j) {
    double df[D], r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = xi[d] - a.pos[j * D + d]; r2 += df[d] * df[d]; }
    if (r2 > a.skip2d) return;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    const double v = j < a.n ? (double)a.vol[j] : ta.wall_volume;
#pragma unroll
    for (int d = 0; d < D; ++d) gr[d].add((2.0 * v) * (df[d] * g));
  });
  double gv[D], sq = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) { gv[d] = -hi(gr[d]); sq += gv[d] * gv[d]; }
  double fac = 1.0;
  if (ta.slope > 0.0) {
    const double x = ta.slope * (ta.h * ta.h * sq);
    fac = x <= 1.0 ? x : 1.0;
  }
  double sh[D], m2 = 0.0;
  bool keep = true;
  if (ta.scope == 1) keep = c[i] >= ta.c_thr;
  if (ta.scope == 2) keep = sqrt(sq) * ta.h <= ta.g_tol;
#pragma unroll
  for (int d = 0; d < D; ++d) { sh[d] = keep ? ta.coef * fac * gv[d] : 0.0; m2 += sh[d] * sh[d]; }
  double mag = sqrt(m2);
  if (ta.cap >= 0.0 && mag > ta.cap) {         // cap < 0: none; 0 clips every shift to zero
    const double sc = ta.cap / (mag > 1.0e-30 ? mag : 1.0e-30);
#pragma unroll
    for (int d = 0; d < D; ++d) sh[d] = sh[d] * sc;
    atomicAdd(&stats[0], 1ull);
    m2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) m2 += sh[d] * sh[d];
    mag = sqrt(m2);
  }
#pragma unroll
  for (int d = 0; d < D; ++d) shift[i * D + d] = sh[d];
  atomicMax(&stats[1], ord_enc(mag));
}

template <int D>
__global__ void k_disp(const double* __restrict__ pos, const double* __restrict__ anchor, int64_t n,
                       unsigned long long* __restrict__ stats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double e = pos[i * D + d] - anchor[i * D + d];
    d2 += e * e;
  }
  atomicMax(&stats[2], ord_enc(sqrt(d2)));
}

template <int D, typename S>
int sim_tvc(tsph_sim* sim, const tsph_tvc_params* prm, double* info, hipStream_t stream) {
  const int64_t n = sim->n;
  // the gradient runs over the step's own lists, as the reference passes the
  // step's pairs; only the bulk scope's concentration needs the pairs within
  // 2h of the current positions
  bool rebuilt = false;
  if (prm->scope == 1) TSPH_TRY((rebuild_keep_volumes<D, S>(sim, stream, &rebuilt)));
  const size_t cb = (sizeof(double) * (size_t)(n + 1) + 255) / 256 * 256;
  const size_t sb = (sizeof(double) * (size_t)n * D + 255) / 256 * 256;
  TSPH_TRY(sim->shift_buf.ensure(2 * cb + sb + 256 + 256));
  char* base = sim->shift_buf.as<char>();
  double* c = (double*)base;
  double* sorted = (double*)(base + cb);
  double* shift = (double*)(base + 2 * cb);
  unsigned long long* st = (unsigned long long*)(base + 2 * cb + sb);
  unsigned long long* cross = st + 4;
  WcsphArgs<S> a = sim_args<D, S>(sim);
  const NbSlots nbr = sim_slots(sim);
  TvcArgs ta{};
  ta.h = sim->prm.op.h;
  ta.coef = prm->alpha * ta.h * ta.h;
  ta.slope = prm->limiter_slope;
  ta.cap = prm->max_shift;
  ta.wall_volume = prm->wall_volume;
  ta.scope = prm->scope;
  ta.g_tol = prm->bulk_g;
  if (prm->scope == 1) {
    k_shift_conc<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, nbr, c);
    TSPH_CHECK_LAUNCH();
    double c_ref = 0.0;
    TSPH_TRY((percentile_of<S>(sim, c, sorted, n, prm->percentile, &c_ref, stream)));
    ta.c_thr = prm->bulk_c * c_ref;
  }
  TSPH_CHECK_HIP(hipMemsetAsync(st, 0, 3 * sizeof(unsigned long long), stream));
  k_tvc_vec<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, nbr, c, ta, shift, st);
  k_shift_move<D><<<grid_for(n), kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), shift,
                                                      sim->anchor.as<double>(), n, st);
  TSPH_CHECK_LAUNCH();
  TSPH_TRY(tsph_apply_boundary(sim->pos[sim->cur].as<double>(), sim->vel[sim->cur].as<double>(), n, D,
                               sim->prm.lo, sim->prm.hi, sim->prm.tol, cross, stream));
  TSPH_CHECK_HIP(hipMemsetAsync(st + 2, 0, sizeof(unsigned long long), stream));
  k_disp<D><<<grid_for(n), kBlock, 0, stream>>>(sim->pos[sim->cur].as<double>(), sim->anchor.as<double>(), n, st);
  TSPH_CHECK_LAUNCH();
  unsigned long long h[5];
  TSPH_CHECK_HIP(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  sim->rebuild_pending = sim->rebuild_forced || (n > 0 && ord_dec_host(h[2]) > 0.5 * sim->prm.skin);
  TSPH_TRY(clear_stale_unless_pending(sim, stream));
  sim->escaped_extra += (int64_t)h[4];
  if (info) {
    info[0] = n > 0 ? ord_dec_host(h[1]) : 0.0;
    info[1] = n > 0 ? (double)h[0] / (double)n : 0.0;
    info[2] = (double)h[4];
  }
  return TSPH_OK;
}

}  // namespace tsph

extern "C" int tsph_sim_tvc(tsph_sim* sim, const tsph_tvc_params* prm, double* info, hipStream_t stream) {
  TSPH_REQUIRE(sim && prm, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(sim->dim == 2, TSPH_ERR_ARG, "the transport-velocity correction runs on the 2-D stepper");
  TSPH_REQUIRE(sim->steps_done > 0, TSPH_ERR_ARG, "the correction needs the volumes of an evaluated step");
  TSPH_REQUIRE(prm->scope >= 0 && prm->scope <= 2, TSPH_ERR_ARG, "scope must be 0 (all), 1 (bulk) or 2 (bulk_g)");
  TSPH_REQUIRE(prm->limiter_slope >= 0.0 && prm->wall_volume >= 0.0, TSPH_ERR_ARG,
               "limiter slope and wall volume must be non-negative");
  TSPH_REQUIRE(prm->scope != 2 || prm->bulk_g >= 0.0, TSPH_ERR_ARG, "scope_tol must be non-negative");
  TSPH_REQUIRE(sim->n < (1ll << 31) - 1, TSPH_ERR_ARG, "too many rows for the concentration sort");
  if (sim->prec == TSPH_F64) return sim_tvc<2, double>(sim, prm, info, stream);
  return sim_tvc<2, float>(sim, prm, info, stream);
}

extern "C" int tsph_sim_shift(tsph_sim* sim, const tsph_shift_params* prm, double* info, hipStream_t stream) {
  TSPH_REQUIRE(sim && prm, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(sim->dim == 2, TSPH_ERR_ARG, "particle shifting runs on the 2-D stepper");
  TSPH_REQUIRE(sim->steps_done > 0, TSPH_ERR_ARG, "shifting needs the volumes of an evaluated step");
  TSPH_REQUIRE(prm->dt >= 0.0 && prm->limit >= 0.0 && prm->strength >= 0.0, TSPH_ERR_ARG,
               "dt, strength and limit must be non-negative");
  TSPH_REQUIRE(sim->n < (1ll << 31) - 1, TSPH_ERR_ARG, "too many rows for the concentration sort");
  if (sim->prec == TSPH_F64) return sim_shift<2, double>(sim, prm, info, stream);
  return sim_shift<2, float>(sim, prm, info, stream);
}
