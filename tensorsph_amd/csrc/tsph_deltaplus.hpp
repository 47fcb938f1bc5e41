// δ⁺-SPH step: per-simulation parameters and the argument block of the four
// sweeps (tsph_deltaplus.hip), shared with the batch engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsph {

// Parameters of one simulation (delta_plus.py:105-130): h = 2 dx,
// rv = 8 nu + alpha h c0 (radial viscosity), diff_scale = 2 delta h c0;
// derived: 1 / h, the Wendland C2 normalisation sigma = 7 / (4 pi h^2) and the
// gradient factor gscale = -5 sigma / h^2 (grad W = gscale (1 - q/2)^3 (x_i - x_j)).
struct DpSim {
  double dx, h, c0, dt, rho0, g, rv, diff_scale, ref_speed;
  double inv_h, sigma, gscale;
};

struct DpArgs {
  int64_t nf, nw;                  // fluid rows, wall stations (all simulations)
  DpSim sim0;                      // used when sims == NULL (one simulation)
  const DpSim* sims;               // per-simulation table (device) or NULL
  const int32_t* fsim;             // simulation of each fluid row, NULL: all 0
  const int32_t* wsim;             // simulation of each wall station, NULL: all 0
  const unsigned char* active;     // per simulation: 0 skips it this step, NULL: all active
  int* status;                     // per simulation: bit 0 wall density <= 0, bit 1 bad density
  const double *x, *v, *rho, *m;   // fluid state (FP64, (nf, 2) / (nf))
  double* vol;                     // per fluid row: m / rho, set by sweep 0
  // Neighbour lists: CSR (off / idx arguments of the sweeps) when fcnt is
  // NULL; else the batch engine's fixed slots, entry k of row r at
  // (r / 64) * 64 * K + 64 k + r % 64 with fcnt / wcnt holding the counts.
  const int32_t *fcnt, *wcnt;
  int64_t kf, kw;
  const double *xw, *vw, *nw_;     // wall stations: position, velocity, unit normal
  double *wall_p, *wall_rho, *reaction;
  double *acc, *vnew, *lam, *inv, *gr, *rho_new, *normal, *strength, *shift, *x_new;
  unsigned char* surface;
  // launch shape of the four pair sweeps: rows per workgroup, and whether
  // consecutive workgroups take contiguous chunks of rows per XCD (1) or
  // round-robin (0); results do not depend on either
  int block = 256;
  int xcd = 1;
};

int dp_queue_step(const DpArgs& a, const int64_t* off, const int32_t* idx, const int64_t* woff,
                  const int32_t* widx, hipStream_t stream);
DpSim dp_sim_params(double dx, double c0, double ref_speed, double dt, double rho0, double g,
                    double nu, double diffusion, double alpha);

}  // namespace tsph
