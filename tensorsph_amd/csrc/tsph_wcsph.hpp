// Argument blocks and device helpers of the WCSPH sweeps (shared by the
// operator entry points and the resident stepper).
#pragma once
#include <string.h>

#include "tsph_common.hpp"
#include "../../include/tsph.h"

namespace tsph {

template <typename S> struct WcsphArgs {
  KernelConst<S> k;
  KernelConst<double> kd;   // the density sweep evaluates W in FP64 in both precisions
  S skip2;            // squared support with a relative margin: beyond it W = grad W = 0
  double skip2d;
  S rho0, bulk, gamma, inv_gamma, ratio_floor, rho_floor;
  double rho0d;       // rho0 in FP64: the EOS forms rho / rho0 - 1 there
  double bulkd, gammad, inv_gammad, ratio_floord;   // the EOS constants in FP64
  S gravity, nu2, friction, eta2, dt, diff_scale, eig_floor;
  int64_t n, nw;
  const double* pos;  // (n + nw, D)
  const double* vel;  // (n + nw, D)
  const S* mass;      // (n + nw)
  const S* mass_lo;   // FP32 resident steps: the part of each mass that FP32 dropped (or NULL)
  const double* wall_accel;
  S* rho;             // (n + nw): fluid rows from sweep A, wall rows from sweep W
  S* p;
  S* pr2;             // p / rho^2
  S* vol;             // m / rho
  S* rho_lo;          // optional: low part of the stored density (FP32 mode)
  Vec4<S>* aux;       // FP32 slot sweeps: (m, q_hi, m / rho, q_lo) per row for the force sweep,
                      // q = p / rho^2 = q_hi + q_lo to about 2^-48
  S* p_lo;            // FP32 resident steps: the part of p that the stored FP32 p dropped
  S* gsum;            // FP32 resident steps: (n, D) sum_j m_j grad W_ij of the density sweep (FP64 sum)
  // opt-in branches of the 2-D operator rhs (unused by the resident steppers)
  double wrap;        // > 0: x separations taken modulo wrap (x-periodic strip)
  int shepard;        // density * lattice_sum / (sum_j W_ij + W(0))
  double lattice_sum;
  const S* dens_off;  // optional additive density offset of the fluid rows
  const S* nu_field;  // optional per-particle viscosity of the fluid rows
};

// Component d of a pair separation: x is reduced to the nearest image when a
// strip width is set (round half to even, as torch.round does).
template <typename S>
TSPH_D double sep(const WcsphArgs<S>& a, double v, int d) {
  return (d == 0 && a.wrap > 0.0) ? v - a.wrap * rint(v / a.wrap) : v;
}

template <typename S>
inline WcsphArgs<S> make_wcsph_args(const tsph_wcsph_params& p, int dim) {
  WcsphArgs<S> a{};
  a.k = make_cubic<S>(p.h, dim);
  a.kd = make_cubic<double>(p.h, dim);
  a.skip2 = S(4.0 * p.h * p.h * (1.0 + 1e-5));
  a.skip2d = 4.0 * p.h * p.h * (1.0 + 1e-5);
  a.rho0 = S(p.rho0);
  a.bulk = S(p.rho0 * (p.sound_speed * p.sound_speed) / p.gamma);
  a.gamma = S(p.gamma);
  a.inv_gamma = S(1.0 / p.gamma);
  a.ratio_floor = S(0.01);
  a.rho_floor = S(0.1 * p.rho0);
  a.rho0d = p.rho0;
  a.bulkd = p.rho0 * (p.sound_speed * p.sound_speed) / p.gamma;
  a.gammad = p.gamma;
  a.inv_gammad = 1.0 / p.gamma;
  a.ratio_floord = 0.01;
  a.gravity = S(p.gravity);
  a.nu2 = S(2.0 * p.viscosity);
  // with walls present and no friction, fluid-wall pairs carry no viscosity
  a.friction = S(p.n_wall > 0 ? p.wall_friction : 1.0);
  a.eta2 = S(0.01 * p.h * p.h);
  a.dt = S(p.dt);
  a.diff_scale = S(p.diffusion_scale);
  a.eig_floor = S(p.eigenvalue_floor);
  a.n = p.n_fluid;
  a.nw = p.n_wall;
  a.wrap = p.wrap_width;
  a.shepard = p.density_renormalization;
  a.lattice_sum = p.lattice_kernel_sum;
  return a;
}

// Order-preserving map of a double onto uint64, so atomicMin/Max on the
// integers orders the doubles (and so is independent of arrival order).
TSPH_D unsigned long long ord_enc(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double ord_dec_host(unsigned long long u) {
  unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  double x;
  memcpy(&x, &b, sizeof(x));
  return x;
}

// Meter slots (per-block partial sums, then atomics on ordered keys).
enum MeterSum { MS_MASS = 0, MS_MX, MS_MY, MS_MZ, MS_KE, MS_RHO, MS_PE, MS_NSUM };
// per-row maxima (reset at every metered step), the per-step displacement and
// the density and speed maxima of every step, from which the resident stepper
// keeps the run's peaks on the host (droplet_impact.py:1697-1705)
enum MeterMax { MM_RHO_MAX = 0, MM_P_MAX, MM_SPEED, MM_ACC, MM_Y_MAX, MM_X_MAX, MM_DISP,
                MM_PEAK_RHO, MM_PEAK_SPEED, MM_NMAX };
enum MeterMin { MN_RHO_MIN = 0, MN_P_MIN, MN_Y_MIN, MN_NMIN };

struct StepArgs {
  double dt;
  double* pos_out;                 // (n + nw, D); wall rows are kept from the input
  double* vel_out;
  double lo[3], hi[3], tol;        // apply_boundary box (droplet_impact.py:1293-1312)
  const double* anchor;            // positions at the last neighbour-list build
  unsigned long long* maxes;       // MM_NMAX ordered keys (device)
  double* block_sums;              // gridDim.x * MS_NSUM partials, or NULL
  // gridDim.x records of 4 keys (displacement, speed, density maxima as
  // ordered keys, boundary hits), or NULL: the sweep adds them into maxes and
  // escaped with atomics
  unsigned long long* block_max;
  unsigned long long* escaped;     // boundary crossings
  int meters;                      // 1: also fill block_sums and the meter maxima
  double gravity;                  // for the potential energy meter
  // clamp_out_of_obstacle (kleefsman_obstacle_3d.py:208-266), dim 3
  int obs_clamp;
  double obs_lo[3], obs_hi[3], obs_eps;
  double* clamp_imp;               // 3 doubles, cumulative m (v_before - v_after)
  unsigned long long* clamp_cnt;
};

// clamp_out_of_obstacle for one particle after the tank clamp: a particle
// strictly inside the box leaves through the nearest face (first of -x, +x,
// -y, +y, -z, +z on a tie) to that face -/+ eps and loses that velocity
// component, whose momentum goes to the impulse accumulator.
// Returns 1 when the particle was moved; its impulse component goes to imp[d]
// (imp may be NULL).
TSPH_D int box_clamp(double* xn, double* vn, double m, const double* lo, const double* hi, double eps,
                     double* imp) {
  bool inside = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) inside = inside && xn[d] > lo[d] && xn[d] < hi[d];
  if (!inside) return 0;
  int face = 0;
  double best = xn[0] - lo[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) {
    const int d = k >> 1;
    const double c = (k & 1) ? hi[d] - xn[d] : xn[d] - lo[d];
    if (c < best) { best = c; face = k; }
  }
  const int d = face >> 1;
  xn[d] = (face & 1) ? hi[d] + eps : lo[d] - eps;
  if (imp) atomicAdd(&imp[d], m * vn[d]);
  vn[d] = 0.0;
  return 1;
}

template <int D>
TSPH_D void obstacle_clamp(double* xn, double* vn, double m, const StepArgs& st) {
  if constexpr (D == 3) {
    if (!st.obs_clamp) return;
    if (box_clamp(xn, vn, m, st.obs_lo, st.obs_hi, st.obs_eps, st.clamp_imp)) atomicAdd(st.clamp_cnt, 1ull);
  }
}

// Regularised inverse f(M) g, f(l) = 1 / max(l, floor_frac * max(trace/D, 1e-30)),
// by cyclic Jacobi on the symmetric matrix built from M's lower triangle (as
// torch.linalg.eigh(UPLO='L') reads it).
template <int D>
TSPH_D void sym_eig(const double* m_in, double* w, double* v) {
  double a[D * D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) a[r * D + c] = r >= c ? m_in[r * D + c] : m_in[c * D + r];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) v[r * D + c] = r == c ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    double offn = 0.0;
#pragma unroll
    for (int p = 0; p < D; ++p)
#pragma unroll
      for (int q = p + 1; q < D; ++q) offn += a[p * D + q] * a[p * D + q];
    if (offn < 1e-300) break;
#pragma unroll
    for (int p = 0; p < D; ++p)
#pragma unroll
      for (int q = p + 1; q < D; ++q) {
        const double apq = a[p * D + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * D + q] - a[p * D + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const double akp = a[k * D + p], akq = a[k * D + q];
          a[k * D + p] = c * akp - s * akq;
          a[k * D + q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const double apk = a[p * D + k], aqk = a[q * D + k];
          a[p * D + k] = c * apk - s * aqk;
          a[q * D + k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const double vkp = v[k * D + p], vkq = v[k * D + q];
          v[k * D + p] = c * vkp - s * vkq;
          v[k * D + q] = s * vkp + c * vkq;
        }
      }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) w[k] = a[k * D + k];
}

template <int D>
TSPH_D void regularised_solve(const double* m, const double* g, double floor_frac, double* out) {
  double w[D], v[D * D];
  sym_eig<D>(m, w, v);
  double tr = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) tr += w[k];
  double mean = tr / D;
  mean = mean > 1e-30 ? mean : 1e-30;
  const double fl = floor_frac * mean;
  double proj[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < D; ++r) s += v[r * D + k] * g[r];
    proj[k] = s / (w[k] > fl ? w[k] : fl);
  }
#pragma unroll
  for (int r = 0; r < D; ++r) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) s += v[r * D + k] * proj[k];
    out[r] = s;
  }
}

}  // namespace tsph
