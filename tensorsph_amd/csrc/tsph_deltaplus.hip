// δ⁺-SPH step (delta_plus.py:105-229, Sun et al. 2017) as per-target gathers.
//
// One step is four sweeps over CSR neighbour lists (fluid targets see fluid
// and wall neighbours, wall targets see fluid neighbours); every pair is
// re-tested against the strict support d^2 < (2h)^2 of the reference
// (delta_plus.py:164) and its kernel values recomputed, so the lists may carry
// a Verlet skin:
//   A  walls  : ghost pressure / density (:172-177), then the wall reaction (:193-197)
//   B  fluid  : pressure + radial artificial viscosity + wall traction, new
//               velocity, first moment M, its regularised inverse and lambda_min
//               (:180-199, _fluid_geometry :34-49), the renormalised density
//               gradient (diffusion)
//   C  fluid  : continuity with slip ghosts + Antuono diffusion (:201-211), the
//               lambda gradient, normal and the empty-cap cover test
//               (_boundary_band :52-69)
//   D  fluid  : band, Sun shift with symmetric density denominator (:213-224),
//               _regularize (:72-76), _wall_tangent (:79-102), new positions.
// Sums run over each target's own list in ascending neighbour order, in FP64.
// Parameters are read per simulation through a sim-id table, so one launch can
// advance many simulations packed into one buffer (the batch engine).
#include <hip/hip_runtime.h>
#include <math.h>

#include "tsph_common.hpp"
#include "tsph_deltaplus.hpp"
#include "tsph_internal.hpp"

namespace tsph {

// Wendland C2 value W and gradient factor s (grad_i W = s (x_i - x_j)) at
// squared separation d2.  sigma dW/dq / (h r) with dW/dq = -5 q (1 - q/2)^3
// reduces to gscale (1 - q/2)^3 for r > 1e-12 (no division per pair); at
// smaller r the reference's clamp of r to 1e-12 is kept.
TSPH_D void dp_kernel(double d2, const DpSim& P, double& w, double& s) {
  const double r = sqrt(d2);
  const double q = r * P.inv_h;
  double t = 1.0 - 0.5 * q;
  t = t < 0.0 ? 0.0 : t;
  const double t2 = t * t, t3 = t2 * t;
  w = P.sigma * (t2 * t2) * (2.0 * q + 1.0);
  s = r > 1.0e-12 ? P.gscale * t3 : P.sigma * (-5.0 * q * t3) / (P.h * 1.0e-12);
}

// separation x_t - x_o and the strict support test, unfused like the reference
TSPH_D bool dp_sep(double xt, double yt, double xo, double yo, double cut2, double& sx, double& sy,
                   double& d2) {
#pragma clang fp contract(off)
  sx = xt - xo;
  sy = yt - yo;
  d2 = sx * sx + sy * sy;
  return d2 < cut2;
}

TSPH_D int dp_sim(const int32_t* ids, int64_t k) { return ids ? ids[k] : 0; }

// Entries of target row t in idx: CSR (entry q at e0 + q), or the batch
// engine's fixed slots interleaved by wave in chunks of four (entry q at
// e0 + 256 (q / 4) + q % 4, e0 = (t / 64) 64 K + 4 (t % 64)).
struct DpRange {
  int64_t e0, n;
  bool slots;
  TSPH_D int64_t at(int64_t q) const { return slots ? e0 + ((q >> 2) << 8) + (q & 3) : e0 + q; }
};
TSPH_D DpRange dp_range(const int64_t* off, const int32_t* cnt, int64_t K, int64_t t) {
  if (cnt) return {(t >> 6) * 64 * K + 4 * (t & 63), (int64_t)cnt[t], true};
  return {off[t], off[t + 1] - off[t], false};
}

// Symmetric 2x2 eigen-decomposition, ascending eigenvalues (one Jacobi rotation
// annihilates the off-diagonal of a 2x2 matrix).
TSPH_D void dp_eig2(const double* m, double* w, double* v) {
  double a00 = m[0], a11 = m[3], a01 = m[2];   // lower triangle, as eigh reads it
  v[0] = 1.0; v[1] = 0.0; v[2] = 0.0; v[3] = 1.0;
  if (a01 != 0.0) {
    const double theta = (a11 - a00) / (2.0 * a01);
    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    a00 = a00 - t * a01;
    a11 = a11 + t * a01;
    v[0] = c; v[1] = s; v[2] = -s; v[3] = c;    // columns are eigenvectors
  }
  w[0] = a00; w[1] = a11;
  if (w[1] < w[0]) {
    double tmp = w[0]; w[0] = w[1]; w[1] = tmp;
    tmp = v[0]; v[0] = v[1]; v[1] = tmp;
    tmp = v[2]; v[2] = v[3]; v[3] = tmp;
  }
}

// Marrone empty-cap test of one neighbour ray (delta_plus.py:58-63)
TSPH_D bool dp_occupied(double rx, double ry, double nx, double ny, double h) {
#pragma clang fp contract(off)
  const double axial = rx * nx + ry * ny;
  const double tangent = fabs(rx * ny - ry * nx);
  const double ox = rx - h * nx, oy = ry - h * ny;
  if (rx * rx + ry * ry >= 2.0 * h * h) return ox * ox + oy * oy < h * h;
  return fabs(axial - h) + tangent < h;
}

TSPH_D double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Wall traction of one fluid-wall pair on the fluid particle (delta_plus.py:186-194).
TSPH_D void dp_wall_force(const DpSim& P, double gx, double gy, double nx, double ny, double vfx, double vfy,
                          double vwx, double vwy, double pf, double pw, double rhof, double volf,
                          double& fx, double& fy) {
  const double sg = sgn(gx * nx + gy * ny);
  const double tx = nx * sg, ty = ny * sg;
  const double closing = (vfx - vwx) * tx + (vfy - vwy) * ty;
  const double gt = gx * tx + gy * ty;
  const double twp = pf + pw + 2.0 * rhof * P.c0 * closing;
  const double c = -(volf * (P.dx * P.dx) * twp);
  fx = c * (gt * tx);
  fy = c * (gt * ty);
}

// ---------------------------------------------------------------- sweep 0 (per-particle volume)
// m / rho of every fluid row, read by the pair loops of sweeps A-C instead of
// a division per pair (the same IEEE quotient)
// the row block of this workgroup: XCD-chunked (consecutive blocks of one
// XCD take neighbouring rows, which then share its L2) or round-robin
TSPH_D int64_t dp_block(const DpArgs& a) {
  return a.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
}

__global__ void __launch_bounds__(kBlock) k_dp_prep(DpArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nf) return;
  const int s = dp_sim(a.fsim, t);
  if (s < 0) return;                          // a hole of the batch layout
  if (a.active && !a.active[s]) return;
  a.vol[t] = a.m[t] / a.rho[t];
}

// ---------------------------------------------------------------- sweep A (walls)
__global__ void __launch_bounds__(kBlock) k_dp_wall(DpArgs a, const int64_t* __restrict__ woff,
                                                    const int32_t* __restrict__ widx) {
  const int64_t w = dp_block(a) * blockDim.x + threadIdx.x;
  if (w >= a.nw) return;
  const int s = dp_sim(a.wsim, w);
  if (s < 0) return;                          // a hole of the batch layout
  if (a.active && !a.active[s]) return;
  const DpSim P = a.sims ? a.sims[s] : a.sim0;
  const DpRange wr = dp_range(woff, a.wcnt, a.kw, w);
  const double cut2 = (2.0 * P.h) * (2.0 * P.h);
  const double xw = a.xw[2 * w], yw = a.xw[2 * w + 1];
  double psum = 0.0, wsum = 0.0;
  int cnt = 0;
  for (int64_t q_ = 0; q_ < wr.n; ++q_) {
    const int64_t e = wr.at(q_);
    const int64_t f = widx[e];
    double sx, sy, d2;
    if (!dp_sep(a.x[2 * f], a.x[2 * f + 1], xw, yw, cut2, sx, sy, d2)) continue;   // x_f - x_w
    double wt, sc;
    dp_kernel(d2, P, wt, sc);
    const double rf = a.rho[f];
    const double pf = (P.c0 * P.c0) * (rf - P.rho0);
    psum += wt * (pf + rf * P.g * sy);
    wsum += wt;
    ++cnt;
  }
  const double pw = psum / (wsum > 2.2250738585072014e-308 ? wsum : 2.2250738585072014e-308);
  const double rw = P.rho0 + pw / (P.c0 * P.c0);
  a.wall_p[w] = pw;
  a.wall_rho[w] = rw;
  if (cnt > 0 && rw <= 0.0) atomicOr(&a.status[s], 1);
  // reaction: minus the traction this station exerts on each fluid neighbour
  const double nx = a.nw_[2 * w], ny = a.nw_[2 * w + 1];
  const double vwx = a.vw[2 * w], vwy = a.vw[2 * w + 1];
  double rx = 0.0, ry = 0.0;
  for (int64_t q_ = 0; q_ < wr.n; ++q_) {
    const int64_t e = wr.at(q_);
    const int64_t f = widx[e];
    double sx, sy, d2;
    if (!dp_sep(a.x[2 * f], a.x[2 * f + 1], xw, yw, cut2, sx, sy, d2)) continue;
    double wt, sc;
    dp_kernel(d2, P, wt, sc);
    const double rf = a.rho[f];
    double fx, fy;
    dp_wall_force(P, sx * sc, sy * sc, nx, ny, a.v[2 * f], a.v[2 * f + 1], vwx, vwy, (P.c0 * P.c0) * (rf - P.rho0),
                  pw, rf, a.vol[f], fx, fy);
    rx += -fx;
    ry += -fy;
  }
  a.reaction[2 * w] = rx;
  a.reaction[2 * w + 1] = ry;
}

// ---------------------------------------------------------------- sweep B (forces, geometry)
__global__ void __launch_bounds__(kBlock) k_dp_force(DpArgs a, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ idx) {
  const int64_t t = dp_block(a) * blockDim.x + threadIdx.x;
  if (t >= a.nf) return;
  const int s = dp_sim(a.fsim, t);
  if (s < 0) return;                          // a hole of the batch layout
  if (a.active && !a.active[s]) return;
  const DpSim P = a.sims ? a.sims[s] : a.sim0;
  const DpRange fr = dp_range(off, a.fcnt, a.kf, t);
  const double cut2 = (2.0 * P.h) * (2.0 * P.h);
  const double xt = a.x[2 * t], yt = a.x[2 * t + 1];
  const double vtx = a.v[2 * t], vty = a.v[2 * t + 1];
  const double rt = a.rho[t], mt = a.m[t];
  const double volt = a.vol[t];
  const double pt = (P.c0 * P.c0) * (rt - P.rho0);
  const bool diff = P.diff_scale != 0.0;
  double nx_ = 0.0, ny_ = 0.0;
  double M[4] = {0.0, 0.0, 0.0, 0.0}, gs[2] = {0.0, 0.0};
  for (int64_t q_ = 0; q_ < fr.n; ++q_) {
    const int64_t e = fr.at(q_);
    const int64_t o = idx[e];
    if (o < a.nf) {
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.x[2 * o], a.x[2 * o + 1], cut2, sx, sy, d2)) continue;
      double wt, sc;
      dp_kernel(d2, P, wt, sc);
      const double gx = sx * sc, gy = sy * sc;
      const double ro = a.rho[o];
      const double volo = a.vol[o];
      const double po = (P.c0 * P.c0) * (ro - P.rho0);
      const double c = -(volt * volo * (pt + po));
      double fx = c * gx, fy = c * gy;
      if (P.rv != 0.0) {
        const double approach = (vtx - a.v[2 * o]) * sx + (vty - a.v[2 * o + 1]) * sy;
        const double fl = 1.0e-12 * P.h * P.h;
        const double pi = approach / (d2 > fl ? d2 : fl);
        const double cv = P.rho0 * P.rv * volt * volo * pi;
        fx += cv * gx;
        fy += cv * gy;
      }
      nx_ += fx;
      ny_ += fy;
      M[0] += volo * (-sx * gx); M[1] += volo * (-sx * gy);
      M[2] += volo * (-sy * gx); M[3] += volo * (-sy * gy);
      if (diff) {
        const double df = ro - rt;
        gs[0] += volo * (df * gx);
        gs[1] += volo * (df * gy);
      }
    } else {
      const int64_t w = o - a.nf;
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.xw[2 * w], a.xw[2 * w + 1], cut2, sx, sy, d2)) continue;
      double wt, sc;
      dp_kernel(d2, P, wt, sc);
      double fx, fy;
      dp_wall_force(P, sx * sc, sy * sc, a.nw_[2 * w], a.nw_[2 * w + 1], vtx, vty, a.vw[2 * w],
                    a.vw[2 * w + 1], pt, a.wall_p[w], rt, volt, fx, fy);
      nx_ += fx;
      ny_ += fy;
    }
  }
  const double ax = nx_ / mt + 0.0, ay = ny_ / mt + (-P.g);
  a.acc[2 * t] = ax;
  a.acc[2 * t + 1] = ay;
  a.vnew[2 * t] = vtx + P.dt * ax;
  a.vnew[2 * t + 1] = vty + P.dt * ay;
  double ev[2], V[4];
  dp_eig2(M, ev, V);
  double fl = 0.1 * (0.5 * (ev[0] + ev[1]));
  fl = fl < 1.0e-12 ? 1.0e-12 : fl;
  const double l0 = ev[0] > fl ? ev[0] : fl, l1 = ev[1] > fl ? ev[1] : fl;
  double inv[4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) inv[r * 2 + c] = V[r * 2 + 0] / l0 * V[c * 2 + 0] + V[r * 2 + 1] / l1 * V[c * 2 + 1];
#pragma unroll
  for (int k = 0; k < 4; ++k) a.inv[4 * t + k] = inv[k];
  a.lam[t] = ev[0];
  if (diff) {
    a.gr[2 * t] = inv[0] * gs[0] + inv[1] * gs[1];
    a.gr[2 * t + 1] = inv[2] * gs[0] + inv[3] * gs[1];
  }
}

// ---------------------------------------------------------------- sweep C (continuity, normal)
__global__ void __launch_bounds__(kBlock) k_dp_continuity(DpArgs a, const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ idx) {
  const int64_t t = dp_block(a) * blockDim.x + threadIdx.x;
  if (t >= a.nf) return;
  const int s = dp_sim(a.fsim, t);
  if (s < 0) return;                          // a hole of the batch layout
  if (a.active && !a.active[s]) return;
  const DpSim P = a.sims ? a.sims[s] : a.sim0;
  const DpRange fr = dp_range(off, a.fcnt, a.kf, t);
  const double cut2 = (2.0 * P.h) * (2.0 * P.h);
  const double xt = a.x[2 * t], yt = a.x[2 * t + 1];
  const double ut = a.vnew[2 * t], wt_ = a.vnew[2 * t + 1];
  const double rt = a.rho[t], lt = a.lam[t];
  const bool diff = P.diff_scale != 0.0;
  const double grt0 = diff ? a.gr[2 * t] : 0.0, grt1 = diff ? a.gr[2 * t + 1] : 0.0;
  double div = 0.0, hr = 0.0, gl0 = 0.0, gl1 = 0.0;
  for (int64_t q_ = 0; q_ < fr.n; ++q_) {
    const int64_t e = fr.at(q_);
    const int64_t o = idx[e];
    if (o < a.nf) {
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.x[2 * o], a.x[2 * o + 1], cut2, sx, sy, d2)) continue;
      double wk, sc;
      dp_kernel(d2, P, wk, sc);
      const double gx = sx * sc, gy = sy * sc;
      const double volo = a.vol[o];
      const double dp = (a.vnew[2 * o] - ut) * gx + (a.vnew[2 * o + 1] - wt_) * gy;
      div += volo * dp;
      if (diff) {
        const double proj = 0.5 * ((grt0 + a.gr[2 * o]) * sx + (grt1 + a.gr[2 * o + 1]) * sy);
        const double contrast = (rt - a.rho[o]) - proj;
        const double radial = -(sx * gx + sy * gy);
        const double cp = contrast * radial / (d2 + 1.0e-12);
        hr += -P.diff_scale * volo * cp;
      }
      const double dl = a.lam[o] - lt;
      gl0 += volo * (dl * gx);
      gl1 += volo * (dl * gy);
    } else {
      const int64_t w = o - a.nf;
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.xw[2 * w], a.xw[2 * w + 1], cut2, sx, sy, d2)) continue;
      double wk, sc;
      dp_kernel(d2, P, wk, sc);
      const double nx = a.nw_[2 * w], ny = a.nw_[2 * w + 1];
      const double ns = (a.vw[2 * w] - ut) * nx + (a.vw[2 * w + 1] - wt_) * ny;
      div += 2.0 * (P.dx * P.dx) * ns * (nx * (sx * sc) + ny * (sy * sc));
    }
  }
  double rate = -rt * div;
  rate += hr;
  const double rn = rt + P.dt * rate;
  a.rho_new[t] = rn;
  if (!(isfinite(rn) && rn > 0.0)) atomicOr(&a.status[s], 2);
  const double* iv = a.inv + 4 * t;
  const double g0 = iv[0] * gl0 + iv[1] * gl1, g1 = iv[2] * gl0 + iv[3] * gl1;
  const double mg = sqrt(g0 * g0 + g1 * g1);
  const double mc = mg > 1.0e-12 ? mg : 1.0e-12;
  const double nx = -g0 / mc, ny = -g1 / mc;
  a.normal[2 * t] = nx;
  a.normal[2 * t + 1] = ny;
  a.strength[t] = mg * P.h;
  // the empty-cap test decides only for 0.2 <= lambda < 0.75 (below: surface,
  // at or above: interior), so the second pass runs for those rows alone
  int covered = 0;
  if (lt >= 0.2 && lt < 0.75) {
    for (int64_t q_ = 0; q_ < fr.n; ++q_) {
      const int64_t e = fr.at(q_);
      const int64_t o = idx[e];
      if (o >= a.nf) continue;
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.x[2 * o], a.x[2 * o + 1], cut2, sx, sy, d2)) continue;
      covered += dp_occupied(-sx, -sy, nx, ny, P.h) ? 1 : 0;
    }
  }
  a.surface[t] = (lt < 0.2) || ((lt < 0.75) && covered == 0);
}

// ---------------------------------------------------------------- sweep D (shift, positions)
__global__ void __launch_bounds__(kBlock) k_dp_shift(DpArgs a, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ idx) {
  const int64_t t = dp_block(a) * blockDim.x + threadIdx.x;
  if (t >= a.nf) return;
  const int s = dp_sim(a.fsim, t);
  if (s < 0) return;                          // a hole of the batch layout
  if (a.active && !a.active[s]) return;
  const DpSim P = a.sims ? a.sims[s] : a.sim0;
  const DpRange fr = dp_range(off, a.fcnt, a.kf, t);
  const double cut2 = (2.0 * P.h) * (2.0 * P.h);
  double wref, sref;
  dp_kernel(P.dx * P.dx, P, wref, sref);
  const double inv_wref = 1.0 / wref;
  const double xt = a.x[2 * t], yt = a.x[2 * t + 1];
  const double rt = a.rho[t];
  double nbs = 0.0, rawx = 0.0, rawy = 0.0, nearest = INFINITY;
  int64_t pick = -1;
  for (int64_t q_ = 0; q_ < fr.n; ++q_) {
    const int64_t e = fr.at(q_);
    const int64_t o = idx[e];
    const bool fluid = o < a.nf;
    const int64_t w = o - a.nf;
    double sx, sy, d2;
    const double xo = fluid ? a.x[2 * o] : a.xw[2 * w], yo = fluid ? a.x[2 * o + 1] : a.xw[2 * w + 1];
    if (!dp_sep(xt, yt, xo, yo, cut2, sx, sy, d2)) continue;
    double wk, sc;
    dp_kernel(d2, P, wk, sc);
    const double ratio = wk * inv_wref;
    const double r2 = ratio * ratio;
    const double f = 1.0 + 0.2 * (r2 * r2);
    const double repx = f * (sx * sc), repy = f * (sy * sc);
    if (fluid) {
      nbs += a.surface[o] ? 1.0 : 0.0;
      const double rden = 1.0 / (rt + a.rho[o]);
      const double mo = a.m[o];
      rawx += mo * (repx * rden);
      rawy += mo * (repy * rden);
    } else {
      const double rw = a.wall_rho[w];
      const double wr = rw * (P.dx * P.dx) / (rt + rw);
      rawx += wr * repx;
      rawy += wr * repy;
      if (d2 < nearest) { nearest = d2; pick = w; }
    }
  }
  const double fac = -(P.dt * P.c0 / P.h) * (P.ref_speed / P.c0) * ((2.0 * P.h) * (2.0 * P.h));
  rawx *= fac;
  rawy *= fac;
  const bool band = a.surface[t] || nbs > 0.0;
  const double nx = a.normal[2 * t], ny = a.normal[2 * t + 1];
  const double rn = rawx * nx + rawy * ny;
  const double tx = rawx - rn * nx, ty = rawy - rn * ny;
  const bool resolved = (a.lam[t] >= 0.55) && (a.strength[t] > 1.0e-6);
  double shx, shy;
  if (band) { shx = resolved ? tx : 0.0; shy = resolved ? ty : 0.0; }
  else { shx = rawx; shy = rawy; }
  if (pick >= 0) {
    const double tol = 1.0e-6 * sqrt(shx * shx + shy * shy);
    const double wnx = a.nw_[2 * pick], wny = a.nw_[2 * pick + 1];
    const double sn = shx * wnx + shy * wny;
    const double gx = shx - sn * wnx, gy = shy - sn * wny;
    bool conflict = band && (fabs(gx * nx + gy * ny) > tol);
    for (int64_t q_ = 0; q_ < fr.n; ++q_) {
      const int64_t e = fr.at(q_);
      const int64_t o = idx[e];
      if (o < a.nf) continue;
      const int64_t w = o - a.nf;
      double sx, sy, d2;
      if (!dp_sep(xt, yt, a.xw[2 * w], a.xw[2 * w + 1], cut2, sx, sy, d2)) continue;
      if (d2 != nearest) continue;
      if (fabs(gx * a.nw_[2 * w] + gy * a.nw_[2 * w + 1]) > tol) conflict = true;
    }
    shx = conflict ? 0.0 : gx;
    shy = conflict ? 0.0 : gy;
  }
  a.shift[2 * t] = shx;
  a.shift[2 * t + 1] = shy;
  a.x_new[2 * t] = (xt + P.dt * a.vnew[2 * t]) + shx;
  a.x_new[2 * t + 1] = (yt + P.dt * a.vnew[2 * t + 1]) + shy;
}

int dp_queue_step(const DpArgs& a, const int64_t* off, const int32_t* idx, const int64_t* woff,
                  const int32_t* widx, hipStream_t stream) {
  // chunked block -> XCD mapping: neighbouring rows are processed on one XCD
  // and share its L2 (round-robin blocks made every L2 fetch every neighbour)
  if (a.nf > 0) k_dp_prep<<<(unsigned)((a.nf + kBlock - 1) / kBlock), kBlock, 0, stream>>>(a);
  const int bs = a.block;
  if (a.nw > 0) k_dp_wall<<<grid_xcd(a.nw, bs), bs, 0, stream>>>(a, woff, widx);
  if (a.nf > 0) {
    k_dp_force<<<grid_xcd(a.nf, bs), bs, 0, stream>>>(a, off, idx);
    k_dp_continuity<<<grid_xcd(a.nf, bs), bs, 0, stream>>>(a, off, idx);
    k_dp_shift<<<grid_xcd(a.nf, bs), bs, 0, stream>>>(a, off, idx);
  }
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

DpSim dp_sim_params(double dx, double c0, double ref_speed, double dt, double rho0, double g,
                    double nu, double diffusion, double alpha) {
  DpSim P{};
  P.dx = dx;
  P.h = 2.0 * dx;
  P.c0 = c0;
  P.dt = dt;
  P.rho0 = rho0;
  P.g = g;
  P.rv = 8.0 * nu + alpha * P.h * c0;
  P.diff_scale = 2.0 * diffusion * P.h * c0;
  P.ref_speed = ref_speed;
  P.inv_h = 1.0 / P.h;
  P.sigma = 7.0 / (4.0 * kPi * P.h * P.h);
  P.gscale = -5.0 * P.sigma / (P.h * P.h);
  return P;
}

}  // namespace tsph

using namespace tsph;

extern "C" int tsph_deltaplus_step(tsph_ctx* ctx, const tsph_dp_params* prm, const double* pos,
                                   const double* vel, const double* rho, const double* mass,
                                   const double* wall, const double* wall_vel, const double* wall_normal,
                                   double* pos_out, double* vel_out, double* rho_out, double* acc_out,
                                   double* reaction, double* shift, hipStream_t stream) {
  TSPH_REQUIRE(ctx && prm, TSPH_ERR_ARG, "ctx and params must not be NULL");
  TSPH_REQUIRE(prm->precision == TSPH_F64, TSPH_ERR_ARG, "the delta-plus step runs in FP64");
  const int64_t n = prm->n, nw = prm->n_wall;
  TSPH_REQUIRE(n > 0 && nw >= 0, TSPH_ERR_ARG, "expected a nonempty fluid");
  TSPH_REQUIRE(ctx->nb_targets == n && ctx->wnb_targets == nw, TSPH_ERR_ARG,
               "neighbour lists hold %lld fluid / %lld wall targets, the step has %lld / %lld "
               "(tsph_nlist_from_pairs with all_targets = 0)", (long long)ctx->nb_targets,
               (long long)ctx->wnb_targets, (long long)n, (long long)nw);
  TSPH_REQUIRE(pos && vel && rho && mass && pos_out && vel_out && rho_out && acc_out && shift,
               TSPH_ERR_ARG, "NULL array");
  TSPH_REQUIRE(nw == 0 || (wall && wall_vel && wall_normal && reaction), TSPH_ERR_ARG, "NULL wall array");
  DpSim P = dp_sim_params(prm->dx, prm->sound_speed, prm->reference_speed, prm->dt, prm->rho0, prm->gravity,
                          prm->viscosity, prm->diffusion, prm->alpha);
  // scratch: sims(1) | status | wall_p, wall_rho (nw) | lam, strength, vol (n) | inv (4n) |
  //          gr, normal (2n) | surface
  const size_t need = 256 + 256 + 8 * (2 * nw + 2) + 8 * (3 * n) + 8 * (4 * n) + 8 * (4 * n) + (size_t)n + 256;
  TSPH_TRY(ctx->s_dp.ensure(need));
  char* base = (char*)ctx->s_dp.ptr;
  DpArgs a{};
  a.sim0 = P;
  a.sims = nullptr;
  a.status = (int*)(base + 256);
  double* d = (double*)(base + 512);
  a.wall_p = d; d += nw + 1;
  a.wall_rho = d; d += nw + 1;
  a.lam = d; d += n;
  a.strength = d; d += n;
  a.vol = d; d += n;
  a.inv = d; d += 4 * n;
  a.gr = d; d += 2 * n;
  a.normal = d; d += 2 * n;
  a.surface = (unsigned char*)d;
  a.nf = n; a.nw = nw;
  a.x = pos; a.v = vel; a.rho = rho; a.m = mass;
  a.xw = wall; a.vw = wall_vel; a.nw_ = wall_normal;
  a.x_new = pos_out; a.vnew = vel_out; a.rho_new = rho_out; a.acc = acc_out; a.reaction = reaction;
  a.shift = shift;
  TSPH_CHECK_HIP(hipMemsetAsync(a.status, 0, sizeof(int), stream));
  TSPH_TRY(dp_queue_step(a, ctx->nb_off.as<int64_t>(), ctx->nb_idx.as<int32_t>(), ctx->wnb_off.as<int64_t>(),
                         ctx->wnb_idx.as<int32_t>(), stream));
  int st = 0;
  TSPH_CHECK_HIP(hipMemcpyAsync(&st, a.status, sizeof(int), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  TSPH_REQUIRE(!(st & 1), TSPH_ERR_NUMERIC, "nonpositive extrapolated wall density");
  TSPH_REQUIRE(!(st & 2), TSPH_ERR_NUMERIC, "delta-plus density became nonpositive or nonfinite");
  return TSPH_OK;
}
