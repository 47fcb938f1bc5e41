// 3-D WCSPH operators of tensorsph.threed (threed/operators3.py:51-295) and a
// resident periodic stepper for the triply periodic shear wave
// (threed/solver.py:211-308).
//
// Operator semantics follow ``accelerations`` (operators3.py:212-256): every
// particle (fluid and wall) is a target; density is the plain kernel sum plus
// the self term (no floor unless density diffusion runs); the Tait law has no
// ratio floor; with the Adami wall model, wall rows take the pressure
// extrapolated from fluid neighbours over raw (not minimum-image) separations
// and the Tait-consistent wall density without a floor; the viscous term is
// the Brookshaw Laplacian divided by the symmetrised discrete second moment
// K_ij / 2, which needs a moment sweep over the final densities.  Pair deltas
// are minimum images when a box is given.  Each particle gathers over its own
// neighbour list, so no two threads write one address and no float atomics
// are used.
#include <hip/hip_runtime.h>

#include <mutex>
#include <math.h>

#include <type_traits>

#include "tsph_cells.hpp"
#include "tsph_common.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"

namespace tsph {

template <typename S> struct ThreedArgs {
  KernelConst<S> k;
  KernelConst<double> kd;   // the density sweep evaluates W in FP64 in both precisions
  S skip2;
  double skip2d;
  S rho0, bulk, gamma, inv_gamma, nu, diff_scale, dt;
  double eig_floor;
  double rho0d;
  double box[3];
  double inv_box[3];   // 1 / box: the image shift rint(v / L) is exact with v * (1 / L) for |v - k L| < L / 3
  double grav[3];
  int periodic, renormalise, has_gravity, adami;
  int do_pressure;   // 0: k3_force skips the pressure term (viscous_acceleration alone)
  int64_t n, nf;
  const double* pos;
  const double* vel;
  const S* mass;
  const S* mass_lo;  // resident FP32 stepper: the part of each mass FP32 dropped (NULL elsewhere)
  S* rho;       // summation density (+ diffusion): the returned rho
  S* rho_u;     // density used by the forces (Adami walls: Tait-consistent)
  S* p;
  S* moment;    // K_i / 3 = (1/3) sum_j r^2 lap_ij m_j / rho_ij
  S* gr;        // renormalised density gradient (Antuono diffusion)
  S* rate_out;  // optional: the diffusion rate of k3_eos
  S* rho_lo;    // low part of the summed density (FP32 mode), NULL in FP64
  Vec4<S>* pack;  // optional: (m, rho_u, p / rho_u^2, K / 3) per row for the force sweep
  int order4;     // fourth-order viscous operator (set only with nu != 0)
  S* fourth;      // I4_i = sum_j r^4 lap_ij m_j / rho_ij (order 4)
  S* lap1;        // (n, 3) first pair Laplacian of the velocity, before nu (order 4)
  // resident periodic stepper: a step whose drift went past half the skin
  // sets this word, and the sweeps of that step (and the queued steps after
  // it) do nothing until the host has rebuilt the lists; NULL elsewhere
  const unsigned int* skip;
};

template <typename S> TSPH_D bool skipped(const ThreedArgs<S>& a) { return a.skip && *a.skip; }

// x_i - x_j from FP64 coordinates, minimum image if asked, rounded to S.
template <typename S>
TSPH_D void delta3(const ThreedArgs<S>& a, int64_t j, const double* xi, S* df, S& r2, bool min_image) {
  r2 = S(0);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = xi[d] - a.pos[j * 3 + d];
    if (min_image) v = v - a.box[d] * rint(v * a.inv_box[d]);
    df[d] = (S)v;
    r2 += df[d] * df[d];
  }
}

TSPH_D void load3(const double* p, int64_t i, double* x) {
  x[0] = p[3 * i]; x[1] = p[3 * i + 1]; x[2] = p[3 * i + 2];
}

// ---------------------------------------------------------------- density
// summation_density (operators3.py:70-79): pair sums, then the self term
// With sup_idx the entries within the support are also written, in list
// order, as the support lists of the later sweeps (NbSlots layout of width K).
// MM selects how the masses are read: 0 one gathered value per entry, 1 the
// two gathered parts (mass + mass_lo) of the FP32 stepper, 2 every particle
// has the same mass (m_0 times the kernel sum, no gather).
template <typename S, typename NB, int MM>
__global__ void __launch_bounds__(kBlock) k3_density(ThreedArgs<S> a, NB nb, S* __restrict__ out,
                                                     int32_t* __restrict__ sup_idx, int32_t* __restrict__ sup_cnt,
                                                     int64_t K) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[3];
  load3(a.pos, i, xi);
  Sum<double> acc;
  SlotWriter sup(sup_idx, K, i);
  const bool per = a.periodic;
  nb.template visit_masked<3>(i, true, [&](int64_t j, bool keep) {
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = xi[d] - a.pos[j * 3 + d];
      v = per ? v - a.box[d] * rint(v * a.inv_box[d]) : v;
      r2 += v * v;
    }
    keep = keep && !(r2 > a.skip2d);
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    if constexpr (MM == 2) {
      acc.add_if(keep, w);
    } else {
      const double mj = MM == 1 ? (double)a.mass[j] + (double)a.mass_lo[j] : (double)a.mass[j];
      acc.add_if(keep, mj * w);
    }
    if (sup_idx) sup.push_if(keep, (int32_t)j);
  });
  if (sup_idx) sup_cnt[i] = sup.finish();
  if constexpr (MM == 2) {
    const double m0 = a.mass_lo ? (double)a.mass[0] + (double)a.mass_lo[0] : (double)a.mass[0];
    store_rho<S>(m0 * (hi(acc) + a.kd.w0d), out, a.rho_lo, i);
  } else {
    const double mi = MM == 1 ? (double)a.mass[i] + (double)a.mass_lo[i] : (double)a.mass[i];
    store_rho<S>(hi(acc) + mi * a.kd.w0d, out, a.rho_lo, i);
  }
}

// Queues k3_density with the mass read that fits the stepper.
template <typename S, typename NB>
void launch_density3(const ThreedArgs<S>& a, NB nb, S* out, int32_t* sup_idx, int32_t* sup_cnt, int64_t K,
                     bool uniform_mass, hipStream_t stream) {
  const unsigned g = grid_xcd(a.n);
  if (uniform_mass)
    k3_density<S, NB, 2><<<g, kBlock, 0, stream>>>(a, nb, out, sup_idx, sup_cnt, K);
  else if (a.mass_lo)
    k3_density<S, NB, 1><<<g, kBlock, 0, stream>>>(a, nb, out, sup_idx, sup_cnt, K);
  else
    k3_density<S, NB, 0><<<g, kBlock, 0, stream>>>(a, nb, out, sup_idx, sup_cnt, K);
}

// renormalised density gradient over fluid-fluid pairs (density_diffusion.py:95-142)
template <typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k3_gradient(ThreedArgs<S> a, NB nb, const S* __restrict__ rho) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (i >= a.nf) { a.gr[3 * i] = a.gr[3 * i + 1] = a.gr[3 * i + 2] = S(0); return; }
  double xi[3];
  load3(a.pos, i, xi);
  const S rhoi = rho[i];
  Sum<S> G[3], M[9];
  nb.template visit<3>(i, true, [&](int64_t j) {
    if (j >= a.nf) return;
    S df[3], r2;
    delta3<S>(a, j, xi, df, r2, a.periodic);
    if (r2 > a.skip2) return;
    S w, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
    const S vj = a.mass[j] / rho[j];
    const S diff = rho[j] - rhoi;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      G[r].add((vj * diff) * (df[r] * g));
#pragma unroll
      for (int c = 0; c < 3; ++c) M[r * 3 + c].add(vj * (-(df[r] * (df[c] * g))));
    }
  });
  double m[9], gv[3], out[3], tr = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    gv[r] = (double)G[r].get();
#pragma unroll
    for (int c = 0; c < 3; ++c) m[r * 3 + c] = (double)M[r * 3 + c].get();
    tr += m[r * 3 + r];
  }
  if (tr <= 0.0) { a.gr[3 * i] = a.gr[3 * i + 1] = a.gr[3 * i + 2] = S(0); return; }
  regularised_solve<3>(m, gv, a.eig_floor, out);
#pragma unroll
  for (int d = 0; d < 3; ++d) a.gr[3 * i + d] = (S)out[d];
}

// Optional density diffusion over fluid-fluid pairs (density_diffusion_rate_3d,
// operators3.py:259-295; rho += dt rate, floored at 0.1 rho0 for every row),
// then the Tait pressure without a ratio floor (operators3.py:82-86).
template <typename S, int FORM, typename NB>
__global__ void __launch_bounds__(kBlock) k3_eos(ThreedArgs<S> a, NB nb, const S* __restrict__ rho_in,
                                                 int diffusion) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double rho_d = (double)rho_in[i] + (a.rho_lo ? (double)a.rho_lo[i] : 0.0);
  if (diffusion) {
    Sum<S> rate;
    if (i < a.nf) {
      double xi[3];
      load3(a.pos, i, xi);
      S gi[3] = {S(0), S(0), S(0)};
      if (FORM == 1) { gi[0] = a.gr[3 * i]; gi[1] = a.gr[3 * i + 1]; gi[2] = a.gr[3 * i + 2]; }
      nb.template visit<3>(i, true, [&](int64_t j) {
        if (j >= a.nf) return;
        S df[3], r2;
        delta3<S>(a, j, xi, df, r2, a.periodic);
        if (r2 > a.skip2) return;
        S w, g;
        cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
        S radial = S(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) radial += df[d] * (df[d] * g);
        radial = -radial;
        S contrast = rho_in[i] - rho_in[j];
        if (FORM == 1) {
          S proj = S(0);
#pragma unroll
          for (int d = 0; d < 3; ++d) proj += (gi[d] + a.gr[3 * j + d]) * df[d];
          contrast = contrast - S(0.5) * proj;
        }
        rate.add(-a.diff_scale * (a.mass[j] / rho_in[j]) * (contrast * radial / (r2 + S(1e-12))));
      });
    }
    if (a.rate_out) a.rate_out[i] = rate.get();
    rho_d = rho_d + (double)a.dt * hi(rate);
    const double fl = 0.1 * a.rho0d;
    rho_d = rho_d > fl ? rho_d : fl;
  }
  const S rho = (S)rho_d;
  const S p = tait_hi<S>(rho_d, a.rho0d, a.bulk, a.gamma, S(-1));
  a.rho[i] = rho;
  a.rho_u[i] = rho;
  a.p[i] = p;
  if (a.pack) a.pack[i] = Vec4<S>{a.mass[i], rho, quot<S>(p, rho * rho), S(0)};
}

// Adami wall pressure (wall_pressure_adami, operators3.py:173-209): raw
// separations, p = 0 without a fluid neighbour, and tait_density
// (operators3.py:89-93) without a floor for the wall density.
template <typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k3_adami(ThreedArgs<S> a, NB nb) {
  if (skipped(a)) return;
  const int64_t w = a.nf + row_block() * blockDim.x + threadIdx.x;
  if (w >= a.n) return;
  double xw[3];
  load3(a.pos, w, xw);
  Sum<S> num, den;
  nb.template visit<3>(w, true, [&](int64_t f) {
    if (f >= a.nf) return;
    S df[3], r2;
    delta3<S>(a, f, xw, df, r2, false);      // x_w - x_f
    if (r2 > a.skip2) return;
    S wt, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, wt, g);
    S donor = a.p[f];
    if (a.has_gravity) {
      const S gd = df[0] * (S)a.grav[0] + df[1] * (S)a.grav[1] + df[2] * (S)a.grav[2];
      donor = donor + a.rho[f] * gd;
    }
    num.add(donor * wt);
    den.add(wt);
  });
  const S dn = den.get();
  const S dd = dn > S(1e-30) ? dn : S(1e-30);
  S pw = dn > S(0) ? num.get() / dd : S(0);
  pw = pw > S(0) ? pw : S(0);
  const S rw = a.rho0 * pow_s<S>(S(1) + pw / a.bulk, a.inv_gamma);
  a.p[w] = pw;
  a.rho_u[w] = rw;
  if (a.pack) a.pack[w] = Vec4<S>{a.mass[w], rw, quot<S>(pw, rw * rw), S(0)};
}

// second moment of the Laplacian (viscous_acceleration, operators3.py:137-145)
// and, for the fourth-order operator, the fourth moment I4 (:147-149)
template <typename S, typename NB>
__global__ void __launch_bounds__(kBlock) k3_moment(ThreedArgs<S> a, NB nb) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[3];
  load3(a.pos, i, xi);
  const S ri = a.rho_u[i];
  Sum<S> acc, acc4;
  auto body = [&](auto four) {
    return [&](int64_t j, bool keep) {
      S df[3], r2;
      delta3<S>(a, j, xi, df, r2, a.periodic);
      keep = keep && !(r2 > a.skip2);
      S g, lap;
      cubic_gl_r2<S>(r2, a.k, g, lap);
      const S rij = S(0.5) * (ri + a.rho_u[j]);
      const S t = r2 * lap * quot<S>(a.mass[j], rij);
      acc.add_if(keep, t);
      if constexpr (decltype(four)::value) acc4.add_if(keep, r2 * t);
    };
  };
  if (a.order4) nb.template visit_masked<3>(i, true, body(std::integral_constant<bool, true>{}));
  else nb.template visit_masked<3>(i, true, body(std::integral_constant<bool, false>{}));
  if (a.order4) a.fourth[i] = acc4.get();
  const S K = acc.get() / S(3);
  a.moment[i] = K;
  if (a.pack) a.pack[i] = Vec4<S>{a.mass[i], ri, quot<S>(a.p[i], ri * ri), K};
}

// Pressure acceleration (operators3.py:96-113) + renormalised viscous
// acceleration (:116-154) + gravity on fluid rows.  With KICK the second
// half-kick of the leapfrog (threed/solver.py:299) is fused: vel_out = v + dt/2 a.
// The reference weights the viscous pair term by the mass of the pair member
// with the higher index (in the caller's order), so that order matters: MW 0
// takes the row order, MW 1 maps rows through ``orig``, MW 2 is for equal
// masses, where either member's mass is the same.
template <typename S, bool KICK, typename NB, int MW, bool PK>
__global__ void __launch_bounds__(kBlock) k3_force(ThreedArgs<S> a, NB nb, const int64_t* __restrict__ orig,
                                                   S* __restrict__ acc_out, double half_dt,
                                                   double* __restrict__ vel_out) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const bool visc = a.nu != S(0);
  double xi[3], vi[3] = {0.0, 0.0, 0.0};
  load3(a.pos, i, xi);
  if (visc || KICK) load3(a.vel, i, vi);     // vel may be NULL for the pressure term alone
  const S ri = a.rho_u[i];
  const S pri = a.do_pressure ? a.p[i] / (ri * ri) : S(0);
  const S Ki = (visc && a.renormalise) ? a.moment[i] : S(0);
  const int64_t oi = MW == 1 ? orig[i] : i;
  const S mi = a.mass[i];
  Sum<S> ap[3], av[3];
  // every list entry is evaluated; the terms of those outside the support are
  // left out of the sums.  The pressure / viscous switches are uniform, so
  // each combination gets its own straight-line loop body.
  auto body = [&](auto press, auto vis) {
    return [&](int64_t j, bool keep) {
      S df[3], r2;
      delta3<S>(a, j, xi, df, r2, a.periodic);
      keep = keep && !(r2 > a.skip2);
      S g, lap;
      cubic_gl_r2<S>(r2, a.k, g, lap);
      Vec4<S> q;                                  // m, rho_u, p / rho_u^2, K / 3 of j
      if constexpr (PK) q = a.pack[j];
      else q.y = a.rho_u[j];
      const S rj = q.y;
      if constexpr (decltype(press)::value) {
        const S sym = pri + (PK ? q.z : quot<S>(a.p[j], rj * rj));
        const S cp = -((PK ? q.x : a.mass[j]) * sym);
#pragma unroll
        for (int d = 0; d < 3; ++d) ap[d].add_if(keep, (df[d] * g) * cp);
      }
      if constexpr (decltype(vis)::value) {
        const S rij = S(0.5) * (ri + rj);
        const S mj = PK ? q.x : a.mass[j];
        S mw;
        if constexpr (MW == 2) {
          mw = mj;
        } else {
          const int64_t oj = MW == 1 ? orig[j] : j;
          mw = oj > oi ? mj : mi;
        }
        S wgt = quot<S>(mw * lap, rij);
        if (a.renormalise) {
          S mk = S(0.5) * (Ki + (PK ? q.w : a.moment[j]));
          mk = mk > S(1e-12) ? mk : S(1e-12);
          wgt = wgt * quot<S>(S(2), mk);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) av[d].add_if(keep, (S)(a.vel[3 * j + d] - vi[d]) * wgt);
      }
    };
  };
  using T1 = std::integral_constant<bool, true>;
  using F0 = std::integral_constant<bool, false>;
  // slot lists: two chunks (eight entries) per block; CSR lists: four entries
  constexpr int W = std::is_same<NB, NbSlots>::value ? 2 : 4;
  if (a.do_pressure && visc) nb.template visit_masked<3, W>(i, true, body(T1{}, T1{}));
  else if (a.do_pressure) nb.template visit_masked<3, W>(i, true, body(T1{}, F0{}));
  else if (visc) nb.template visit_masked<3, W>(i, true, body(F0{}, T1{}));
  if (a.order4) {
    // the fourth-order operator finishes in k3_visc4, which needs every row's
    // first Laplacian: store it with the pressure term and stop here
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc_out[3 * i + d] = ap[d].get();
      a.lap1[3 * i + d] = av[d].get();
    }
    return;
  }
  S acc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    acc[d] = ap[d].get() + av[d].get() * a.nu;
    if (a.has_gravity && i < a.nf) acc[d] = acc[d] + (S)a.grav[d];
    acc_out[3 * i + d] = acc[d];
  }
  if (KICK) {
#pragma unroll
    for (int d = 0; d < 3; ++d) vel_out[3 * i + d] = vi[d] + half_dt * (double)acc[d];
  }
}

// Second pass of the fourth-order viscous operator (operators3.py:146-151):
// L4 v_i = L v_i - sum_j w_ij beta_ij (L v_j - L v_i), with w_ij the pair
// weight of k3_force (same member mass, same renormalisation) and beta_ij =
// (I4_i + I4_j) / 240.  acc_out holds the pressure term of k3_force on entry;
// the row's acceleration is that + nu L4 v (+ gravity on fluid rows), and
// with KICK the second half-kick follows as in k3_force.
template <typename S, bool KICK, typename NB, int MW>
__global__ void __launch_bounds__(kBlock) k3_visc4(ThreedArgs<S> a, NB nb, const int64_t* __restrict__ orig,
                                                   S* __restrict__ acc_out, double half_dt,
                                                   double* __restrict__ vel_out) {
  if (skipped(a)) return;
  const int64_t i = row_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[3];
  load3(a.pos, i, xi);
  const S ri = a.rho_u[i];
  const S Ki = a.renormalise ? a.moment[i] : S(0);
  const S Fi = a.fourth[i];
  const int64_t oi = MW == 1 ? orig[i] : i;
  const S mi = a.mass[i];
  S li[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) li[d] = a.lap1[3 * i + d];
  Sum<S> s[3];
  constexpr int W = std::is_same<NB, NbSlots>::value ? 2 : 4;
  nb.template visit_masked<3, W>(i, true, [&](int64_t j, bool keep) {
    S df[3], r2;
    delta3<S>(a, j, xi, df, r2, a.periodic);
    keep = keep && !(r2 > a.skip2);
    S g, lap;
    cubic_gl_r2<S>(r2, a.k, g, lap);
    const S rij = S(0.5) * (ri + a.rho_u[j]);
    const S mj = a.mass[j];
    S mw;
    if constexpr (MW == 2) {
      mw = mj;
    } else {
      const int64_t oj = MW == 1 ? orig[j] : j;
      mw = oj > oi ? mj : mi;
    }
    S wgt = quot<S>(mw * lap, rij);
    if (a.renormalise) {
      S mk = S(0.5) * (Ki + a.moment[j]);
      mk = mk > S(1e-12) ? mk : S(1e-12);
      wgt = wgt * quot<S>(S(2), mk);
    }
    const S wb = wgt * quot<S>(S(0.5) * (Fi + a.fourth[j]), S(120));
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d].add_if(keep, (a.lap1[3 * j + d] - li[d]) * wb);
  });
  S acc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    acc[d] = acc_out[3 * i + d] + (li[d] - s[d].get()) * a.nu;
    if (a.has_gravity && i < a.nf) acc[d] = acc[d] + (S)a.grav[d];
    acc_out[3 * i + d] = acc[d];
  }
  if (KICK) {
#pragma unroll
    for (int d = 0; d < 3; ++d) vel_out[3 * i + d] = a.vel[3 * i + d] + half_dt * (double)acc[d];
  }
}

// the force sweep of one evaluation; the fourth-order operator adds a second
// sweep, which takes over the half-kick
template <typename S, bool KICK, typename NB, int MW, bool PK>
void queue_force(const ThreedArgs<S>& a, NB nb, const int64_t* orig, S* acc_out, double half_dt,
                 double* vel_out, hipStream_t stream) {
  const unsigned g = grid_xcd(a.n);
  if (a.order4) {
    k3_force<S, false, NB, MW, PK><<<g, kBlock, 0, stream>>>(a, nb, orig, acc_out, 0.0, nullptr);
    k3_visc4<S, KICK, NB, MW><<<g, kBlock, 0, stream>>>(a, nb, orig, acc_out, half_dt, vel_out);
  } else {
    k3_force<S, KICK, NB, MW, PK><<<g, kBlock, 0, stream>>>(a, nb, orig, acc_out, half_dt, vel_out);
  }
}

struct ThreedPrm {
  double h, rho0, c0, gamma, nu, diff_scale, eig_floor, dt;
  double box[3], grav[3];
  int periodic, has_grav, renormalise, diff_form, adami, order4;
};

inline ThreedPrm threed_prm(const tsph_threed_params& q) {
  ThreedPrm r{};
  r.h = q.h; r.rho0 = q.rho0; r.c0 = q.sound_speed; r.gamma = q.gamma; r.nu = q.viscosity;
  r.diff_scale = q.diffusion_scale; r.eig_floor = q.eigenvalue_floor; r.dt = q.dt;
  r.periodic = q.periodic; r.has_grav = q.has_gravity; r.renormalise = q.renormalise;
  r.diff_form = q.diffusion_form; r.adami = q.wall_model == 0;
  r.order4 = q.viscous_order == 4;
  for (int d = 0; d < 3; ++d) { r.box[d] = q.periodic ? q.box[d] : 1.0; r.grav[d] = q.has_gravity ? q.gravity[d] : 0.0; }
  return r;
}

template <typename S>
ThreedArgs<S> make_threed_args(const ThreedPrm& q, int64_t n, int64_t nf) {
  ThreedArgs<S> a{};
  a.k = make_cubic<S>(q.h, 3);
  a.kd = make_cubic<double>(q.h, 3);
  a.skip2 = S(4.0 * q.h * q.h * (1.0 + 1e-5));
  a.skip2d = 4.0 * q.h * q.h * (1.0 + 1e-5);
  a.rho0 = S(q.rho0);
  a.rho0d = q.rho0;
  a.bulk = S(q.rho0 * (q.c0 * q.c0) / q.gamma);
  a.gamma = S(q.gamma);
  a.inv_gamma = S(1.0 / q.gamma);
  a.nu = S(q.nu);
  a.diff_scale = S(q.diff_scale);
  a.eig_floor = q.eig_floor;
  a.dt = S(q.dt);
  a.periodic = q.periodic;
  for (int d = 0; d < 3; ++d) { a.box[d] = q.box[d]; a.inv_box[d] = 1.0 / q.box[d]; a.grav[d] = q.grav[d]; }
  a.renormalise = q.renormalise;
  a.has_gravity = q.has_grav;
  a.adami = q.adami;
  a.do_pressure = 1;
  a.order4 = q.order4 && q.nu != 0.0;
  a.n = n;
  a.nf = nf;
  return a;
}

// scratch of one evaluation, n + 1 elements each (gr: 3 (n + 1))
template <typename S> struct ThreedScratch {
  S *rho_raw, *rho, *rho_u, *p, *moment, *gr, *rho_lo;   // rho_lo: NULL in FP64 mode
  Vec4<S>* pack;
  S *fourth, *lap1;   // fourth-order viscous operator only (n + 1, 3 (n + 1)); else NULL
};

// Queues the sweeps of one accelerations() evaluation over a neighbour source
// (CSR lists or slot lists).  uniform_mass: every particle has the same mass.
template <typename S, bool KICK, typename NB>
int threed_sweeps(ThreedArgs<S> a, const ThreedPrm& q, const ThreedScratch<S>& b, NB nb, const int64_t* orig,
                  S* acc_out, double half_dt, double* vel_out, hipStream_t stream, bool uniform_mass = false,
                  NB* support = nullptr) {
  const int64_t n = a.n;
  if (n == 0) return TSPH_OK;
  a.rho = b.rho; a.rho_u = b.rho_u; a.p = b.p; a.moment = b.moment; a.gr = b.gr; a.rho_lo = b.rho_lo;
  a.pack = b.pack; a.fourth = b.fourth; a.lap1 = b.lap1;
  const unsigned g = grid_xcd(n);
  const bool diff = q.diff_scale > 0.0;
  // with support lists (slot lists only) the density sweep writes them and
  // every later sweep reads them instead of the build-radius lists
  if constexpr (std::is_same<NB, NbSlots>::value) {
    if (support) {
      launch_density3<S, NB>(a, nb, diff ? b.rho_raw : b.rho, const_cast<int32_t*>(support->idx),
                             const_cast<int32_t*>(support->cnt), nb.K, uniform_mass, stream);
      nb = *support;
    } else {
      launch_density3<S, NB>(a, nb, diff ? b.rho_raw : b.rho, nullptr, nullptr, 0, uniform_mass, stream);
    }
  } else {
    launch_density3<S, NB>(a, nb, diff ? b.rho_raw : b.rho, nullptr, nullptr, 0, uniform_mass, stream);
  }
  if (diff) {
    if (q.diff_form == 1) {
      k3_gradient<S, NB><<<g, kBlock, 0, stream>>>(a, nb, b.rho_raw);
      k3_eos<S, 1, NB><<<g, kBlock, 0, stream>>>(a, nb, b.rho_raw, 1);
    } else {
      k3_eos<S, 0, NB><<<g, kBlock, 0, stream>>>(a, nb, b.rho_raw, 1);
    }
  } else {
    k3_eos<S, 0, NB><<<g, kBlock, 0, stream>>>(a, nb, b.rho, 0);
  }
  if (a.nf < n && a.adami) k3_adami<S, NB><<<grid_xcd(n - a.nf), kBlock, 0, stream>>>(a, nb);
  if (a.nu != S(0) && (a.renormalise || a.order4)) k3_moment<S, NB><<<g, kBlock, 0, stream>>>(a, nb);
  if (uniform_mass)
    queue_force<S, KICK, NB, 2, true>(a, nb, orig, acc_out, half_dt, vel_out, stream);
  else if (orig)
    queue_force<S, KICK, NB, 1, true>(a, nb, orig, acc_out, half_dt, vel_out, stream);
  else
    queue_force<S, KICK, NB, 0, true>(a, nb, orig, acc_out, half_dt, vel_out, stream);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

template <typename S>
int threed_ctx_scratch(tsph_ctx* ctx, int64_t n, ThreedScratch<S>& b) {
  const size_t s = sizeof(S) * (n + 1);
  TSPH_TRY(ctx->s_rho_raw.ensure(s));
  TSPH_TRY(ctx->s_rho.ensure(s));
  TSPH_TRY(ctx->s_aux.ensure(s));
  TSPH_TRY(ctx->s_p.ensure(s));
  TSPH_TRY(ctx->s_aux2.ensure(s));
  TSPH_TRY(ctx->s_gr.ensure(3 * s));
  TSPH_TRY(ctx->s_aux3.ensure(s));
  TSPH_TRY(ctx->s_pack.ensure(4 * s));
  TSPH_TRY(ctx->s_four.ensure(s));
  TSPH_TRY(ctx->s_lap1.ensure(3 * s));
  b.fourth = ctx->s_four.as<S>(); b.lap1 = ctx->s_lap1.as<S>();
  b.rho_lo = sizeof(S) == 4 ? ctx->s_aux3.as<S>() : nullptr;
  b.pack = ctx->s_pack.as<Vec4<S>>();
  b.rho_raw = ctx->s_rho_raw.as<S>(); b.rho = ctx->s_rho.as<S>(); b.rho_u = ctx->s_aux.as<S>();
  b.p = ctx->s_p.as<S>(); b.moment = ctx->s_aux2.as<S>(); b.gr = ctx->s_gr.as<S>();
  return TSPH_OK;
}

template <typename S>
int threed_accel(tsph_ctx* ctx, const tsph_threed_params* prm, const double* pos, const double* vel,
                 const void* mass, void* acc, void* rho, void* p, hipStream_t stream) {
  const int64_t n = prm->n;
  if (n == 0) return TSPH_OK;
  ThreedScratch<S> b;
  TSPH_TRY(threed_ctx_scratch<S>(ctx, n, b));
  const ThreedPrm q = threed_prm(*prm);
  ThreedArgs<S> a = make_threed_args<S>(q, n, prm->n_fluid);
  a.pos = pos; a.vel = vel; a.mass = (const S*)mass;
  const NbList nb{ctx->nb_off.as<int64_t>(), ctx->nb_idx.as<int32_t>()};
  TSPH_TRY((threed_sweeps<S, false>(a, q, b, nb, nullptr, (S*)acc, 0.0, nullptr, stream)));
  if (rho) TSPH_CHECK_HIP(hipMemcpyAsync(rho, b.rho, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
  if (p) TSPH_CHECK_HIP(hipMemcpyAsync(p, b.p, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
  return TSPH_OK;
}

// One operator of threed.operators3 on its own, with the caller's rho / p.
template <typename S>
int threed_operator(tsph_ctx* ctx, const tsph_threed_params* prm, int op, const double* pos, const double* vel,
                    const void* mass, const void* rho, const void* p, void* out, hipStream_t stream) {
  const int64_t n = prm->n, nf = prm->n_fluid;
  if (n == 0) return TSPH_OK;
  ThreedScratch<S> b;
  TSPH_TRY(threed_ctx_scratch<S>(ctx, n, b));
  ThreedPrm q = threed_prm(*prm);
  ThreedArgs<S> a = make_threed_args<S>(q, n, nf);
  a.pos = pos; a.vel = vel; a.mass = (const S*)mass;
  a.rho = b.rho; a.rho_u = b.rho_u; a.p = b.p; a.moment = b.moment; a.gr = b.gr;
  a.fourth = b.fourth; a.lap1 = b.lap1;
  const NbList nb{ctx->nb_off.as<int64_t>(), ctx->nb_idx.as<int32_t>()};
  const unsigned g = grid_xcd(n);
  S* rho_in = (S*)rho;           // read only
  S* p_in = (S*)p;               // read only
  switch (op) {
    case 0:
      launch_density3<S, NbList>(a, nb, (S*)out, nullptr, nullptr, 0, false, stream);
      break;
    case 1:
      a.rho_u = rho_in; a.p = p_in; a.nu = S(0); a.has_gravity = 0; a.order4 = 0;
      k3_force<S, false, NbList, 0, false><<<g, kBlock, 0, stream>>>(a, nb, nullptr, (S*)out, 0.0, nullptr);
      break;
    case 2:
      a.rho_u = rho_in; a.do_pressure = 0; a.has_gravity = 0;
      if (a.nu != S(0) && (a.renormalise || a.order4)) k3_moment<S, NbList><<<g, kBlock, 0, stream>>>(a, nb);
      queue_force<S, false, NbList, 0, false>(a, nb, nullptr, (S*)out, 0.0, nullptr, stream);
      break;
    case 3:
      if (nf == n) return TSPH_OK;
      TSPH_CHECK_HIP(hipMemcpyAsync(b.p, p_in, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
      a.rho = rho_in;
      k3_adami<S, NbList><<<grid_xcd(n - nf), kBlock, 0, stream>>>(a, nb);
      TSPH_CHECK_LAUNCH();
      TSPH_CHECK_HIP(hipMemcpyAsync(out, b.p + nf, sizeof(S) * (n - nf), hipMemcpyDeviceToDevice, stream));
      break;
    case 4:
      a.rate_out = (S*)out;
      if (q.diff_form == 1) {
        k3_gradient<S, NbList><<<g, kBlock, 0, stream>>>(a, nb, rho_in);
        k3_eos<S, 1, NbList><<<g, kBlock, 0, stream>>>(a, nb, rho_in, 1);
      } else {
        k3_eos<S, 0, NbList><<<g, kBlock, 0, stream>>>(a, nb, rho_in, 1);
      }
      break;
    default:
      TSPH_REQUIRE(false, TSPH_ERR_ARG, "unknown operator %d", op);
  }
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

// ---------------------------------------------------------------- periodic stepper kernels
// flag = 1 when some mass differs from the first one
template <typename S>
__global__ void k3_mass_differs(const S* __restrict__ m, int64_t n, unsigned int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && m[i] != m[0]) atomicOr(flag, 1u);
}

__global__ void k3_iota(int64_t* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// torch.remainder for a positive period: fmod, shifted into [0, L) when negative
TSPH_D double wrap_period(double x, double L) {
  double m = fmod(x, L);
  if (m != 0.0 && m < 0.0) m += L;
  return m;
}

// First half-kick and drift of the leapfrog (threed/solver.py:287-289), then
// max |minimum_image(x - anchor)| for the skin test (:293-295).
template <typename S>
__global__ void __launch_bounds__(kBlock) k3_kick_drift(double* __restrict__ pos, double* __restrict__ vel,
                                                        const S* __restrict__ acc, int64_t n, double dt,
                                                        double b0, double b1, double b2,
                                                        const double* __restrict__ anchor,
                                                        unsigned long long* __restrict__ disp,
                                                        const unsigned int* __restrict__ stale) {
  if (*stale) return;                      // an earlier step awaits its list rebuild
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    const double box[3] = {b0, b1, b2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = vel[3 * i + d] + 0.5 * dt * (double)acc[3 * i + d];
      const double x = wrap_period(pos[3 * i + d] + dt * v, box[d]);
      vel[3 * i + d] = v;
      pos[3 * i + d] = x;
      double m = x - anchor[3 * i + d];
      m = m - box[d] * rint(m / box[d]);
      d2 += m * m;
    }
  }
  double v = sqrt(d2);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  // one atomic per block: per-wave atomics on the one word serialise
  __shared__ double wv[kBlock / 64];
  if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, wv[w]);
    atomicMax(disp, ord_enc(v));
  }
}

// The skin test of the step just drifted (threed/solver.py:293-299): a
// displacement past half the skin marks the lists stale (the step's sweeps
// and every queued step then do nothing); the flag also goes to the host's
// ring entry of the step.  Resets the displacement maximum for the next step.
__global__ void k3_skin_test(unsigned long long* __restrict__ disp, double skin, unsigned int* __restrict__ stale,
                             unsigned int* __restrict__ ring_entry) {
  if (threadIdx.x != 0) return;
  const unsigned long long u = *disp;
  const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  const double d = __longlong_as_double((long long)b);
  if (!*stale && 2.0 * d > skin) *stale = 1u;
  *disp = 0ull;
  *ring_entry = *stale;
}

// A kicked evaluation that was skipped still moves the velocities to the
// other buffer, which the host switches to after every evaluation.
__global__ void k3_carry_vel(const unsigned int* __restrict__ stale, const double* __restrict__ vel,
                             double* __restrict__ vel_out, int64_t n3) {
  if (!*stale) return;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) vel_out[k] = vel[k];
}

// row k's mass from the caller's FP64 masses (row k holds caller row
// orig[k]): FP32 runs as hi + lo (lo at lo_off), FP64 runs as given
template <typename S>
__global__ void k3_set_masses(const double* __restrict__ m, const int64_t* __restrict__ orig, int64_t n,
                              S* __restrict__ mass, int64_t lo_off) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double v = m[orig[k]];
  const S h = (S)v;
  mass[k] = h;
  if (lo_off) mass[lo_off + k] = (S)(v - (double)h);
}

// Re-orders the state by the cell order of the last search; afterwards the
// search's own slot order is the identity.
template <typename S>
__global__ void __launch_bounds__(kBlock) k3_permute(const int64_t* __restrict__ order, int64_t n,
                                                     const double* __restrict__ pos, const double* __restrict__ vel,
                                                     const S* __restrict__ mass, const S* __restrict__ acc,
                                                     const int64_t* __restrict__ orig,
                                                     const int64_t* __restrict__ cells, double* __restrict__ pos2,
                                                     double* __restrict__ vel2, S* __restrict__ mass2,
                                                     S* __restrict__ acc2, int64_t* __restrict__ orig2,
                                                     int64_t* __restrict__ cells2, int64_t lo_off) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t o = order[k];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pos2[3 * k + d] = pos[3 * o + d];
    vel2[3 * k + d] = vel[3 * o + d];
    acc2[3 * k + d] = acc[3 * o + d];
    cells2[3 * k + d] = cells[3 * o + d];
  }
  mass2[k] = mass[o];
  if (lo_off) mass2[lo_off + k] = mass[lo_off + o];   // the low parts (FP32 runs)
  orig2[k] = orig[o];
}

__global__ void k3_order_identity(int64_t* __restrict__ order, int64_t n) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) order[s] = s;
}

// ---------------------------------------------------------------- diagnostics
// conservation.diagnostics (threed/conservation.py:106-121) and the shear-wave
// mode projections (operators3.py:401-408): per-block partial sums in FP64,
// reduced in a fixed order.
enum DiagSlot {
  DG_MASS = 0, DG_PX, DG_PY, DG_PZ, DG_LX, DG_LY, DG_LZ, DG_KE, DG_IE,
  DG_SX, DG_SY, DG_SZ, DG_CX, DG_CY, DG_CZ, DG_MODE1, DG_MODE2, DG_ABSMV, DG_RX, DG_RY, DG_RZ, DG_N
};
static_assert(DG_N == 21, "diagnostic row width is part of the ABI");

template <typename S>
__global__ void __launch_bounds__(kBlock) k3_diag(const double* __restrict__ pos, const double* __restrict__ vel,
                                                  const S* __restrict__ mass, const S* __restrict__ rho, int64_t n,
                                                  double rho0, double c0, double gamma, double b0, double b1,
                                                  double b2, double wavenumber, double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s[DG_N];
#pragma unroll
  for (int k = 0; k < DG_N; ++k) s[k] = 0.0;
  if (i < n) {
    const double m = (double)mass[i];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double u = vel[3 * i], v = vel[3 * i + 1], w = vel[3 * i + 2];
    s[DG_MASS] = m;
    s[DG_PX] = m * u; s[DG_PY] = m * v; s[DG_PZ] = m * w;
    s[DG_LX] = m * (y * w - z * v); s[DG_LY] = m * (z * u - x * w); s[DG_LZ] = m * (x * v - y * u);
    s[DG_KE] = 0.5 * (m * (u * u) + m * (v * v) + m * (w * w));
    const double ratio = (double)rho[i] / rho0;
    const double e = (c0 * c0 / gamma) * ((pow(ratio, gamma - 1.0) - 1.0) / (gamma - 1.0) + 1.0 / ratio - 1.0);
    s[DG_IE] = m * e;
    const double tp = 2.0 * kPi;
    s[DG_SX] = m * sin(tp * x / b0); s[DG_SY] = m * sin(tp * y / b1); s[DG_SZ] = m * sin(tp * z / b2);
    s[DG_CX] = m * cos(tp * x / b0); s[DG_CY] = m * cos(tp * y / b1); s[DG_CZ] = m * cos(tp * z / b2);
    s[DG_MODE1] = u * sin(wavenumber * z);
    s[DG_MODE2] = u * (sin(wavenumber * y) * sin(wavenumber * z));
    s[DG_ABSMV] = m * fabs(u) + m * fabs(v) + m * fabs(w);
    s[DG_RX] = m * x; s[DG_RY] = m * y; s[DG_RZ] = m * z;
  }
  __shared__ double red[DG_N][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < DG_N; ++k) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < DG_N) {
    double v = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) v += red[threadIdx.x][w];
    part[(int64_t)blockIdx.x * DG_N + threadIdx.x] = v;
  }
}

__global__ void __launch_bounds__(256) k3_diag_reduce(const double* __restrict__ part, int64_t nblocks,
                                                      double* __restrict__ out) {
  __shared__ double red[DG_N][256];
  for (int k = 0; k < DG_N; ++k) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 256) v += part[b * DG_N + k];
    red[k][threadIdx.x] = v;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < DG_N; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < DG_N) out[threadIdx.x] = red[threadIdx.x][0];
}

template <typename S>
__global__ void __launch_bounds__(kBlock) k3_unpermute(const int64_t* __restrict__ orig, int64_t n,
                                                       const double* __restrict__ pos,
                                                       const double* __restrict__ vel, const S* __restrict__ rho,
                                                       const S* __restrict__ acc, double* pos_o, double* vel_o,
                                                       S* rho_o, S* acc_o) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t o = orig[k];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (pos_o) pos_o[3 * o + d] = pos[3 * k + d];
    if (vel_o) vel_o[3 * o + d] = vel[3 * k + d];
    if (acc_o) acc_o[3 * o + d] = acc[3 * k + d];
  }
  if (rho_o) rho_o[o] = rho[k];
}

}  // namespace tsph

using namespace tsph;

// ---------------------------------------------------------------- periodic stepper state
struct tsph_psim {
  tsph_threed_params prm{};
  ThreedPrm q{};
  double skin = 0.0;
  int64_t n = 0;
  size_t s = 8;
  tsph_ctx* ctx = nullptr;
  DevBuf pos, vel[2], mass, acc, orig, anchor, disp;
  DevBuf pos2, vel2, mass2, acc2, orig2, cells2, part, diag;
  DevBuf rho_raw, rho, rho_u, p, moment, gr, rho_lo, pack;
  DevBuf cell_R, entries;                    // dense cell table; list entries at the build radius
  DevBuf slot_idx, slot_cnt, slot_max;       // per-row lists in fixed slots (NbSlots layout)
  DevBuf slot_posf;                          // FP32 build positions (list prefilter)
  DevBuf sup_idx, sup_cnt;                   // this evaluation's entries within the support
  DevBuf fourth, lap1;                       // fourth-order viscous operator only
  int64_t slot_K = 0;
  int64_t ncells_dense = 0, cell_ext[3] = {1, 1, 1}, cell_stride[3] = {0, 0, 0};
  unsigned long long* disp_host = nullptr;   // pinned
  // lagged skin test: stale word (device), per-step ring of its value (pinned)
  // and the events that complete each queued step
  static constexpr int kRingP = 4;
  DevBuf stale;
  unsigned int* ring = nullptr;              // pinned, kRingP
  hipEvent_t ev[kRingP] = {};
  int cur = 0;
  int64_t steps = 0, rebuilds = 0;
  bool built = false, have_acc = false;
  bool uniform_mass = false;                 // every particle has the same mass
};

namespace tsph {

template <typename S> ThreedScratch<S> psim_scratch(tsph_psim* ps) {
  ThreedScratch<S> b;
  b.rho_raw = ps->rho_raw.as<S>(); b.rho = ps->rho.as<S>(); b.rho_u = ps->rho_u.as<S>();
  b.p = ps->p.as<S>(); b.moment = ps->moment.as<S>(); b.gr = ps->gr.as<S>();
  b.rho_lo = sizeof(S) == 4 ? ps->rho_lo.as<S>() : nullptr;
  b.pack = ps->pack.as<Vec4<S>>();
  b.fourth = ps->fourth.as<S>(); b.lap1 = ps->lap1.as<S>();
  return b;
}

static NbSlots psim_slots(const tsph_psim* ps) {
  NbSlots l{};
  l.idx = ps->slot_idx.as<int32_t>();
  l.cnt = ps->slot_cnt.as<int32_t>();
  l.K = ps->slot_K;
  l.nf = ps->n;
  return l;
}

static NbCells psim_cells(const tsph_psim* ps) {
  NbCells c{};
  c.R = ps->cell_R.as<int4>();
  c.cells = ps->cells2.as<int64_t>();
  for (int d = 0; d < 3; ++d) { c.ext[d] = ps->cell_ext[d]; c.stride[d] = ps->cell_stride[d]; }
  c.periodic = 1;
  return c;
}

// periodic_pairs at cutoff 2h + skin (threed/solver.py:254-256, 296-298):
// the state is re-ordered by (Morton) cell; a dense table of per-cell row
// ranges drives a walk of the 27 cells around every particle that writes each
// row's list, in the walk's order, into fixed slots interleaved by wave
template <typename S>
int psim_rebuild(tsph_psim* ps, hipStream_t stream) {
  tsph_ctx* ctx = ps->ctx;
  const int64_t n = ps->n;
  const double rc = 2.0 * ps->prm.h + ps->skin;
  TSPH_TRY((run_search<double, 3>(ctx, ps->pos.as<double>(), n, rc, ps->prm.box, stream, true)));
  CellSearch& cs = ctx->search;
  const unsigned g = grid_for(n);
  k3_permute<S><<<g, kBlock, 0, stream>>>(cs.order.as<int64_t>(), n, ps->pos.as<double>(),
                                          ps->vel[ps->cur].as<double>(), ps->mass.as<S>(), ps->acc.as<S>(),
                                          ps->orig.as<int64_t>(), cs.cells.as<int64_t>(), ps->pos2.as<double>(),
                                          ps->vel2.as<double>(), ps->mass2.as<S>(), ps->acc2.as<S>(),
                                          ps->orig2.as<int64_t>(), ps->cells2.as<int64_t>(),
                                          sizeof(S) == 4 ? (int64_t)n : 0);
  k3_order_identity<<<g, kBlock, 0, stream>>>(cs.order.as<int64_t>(), n);
  TSPH_CHECK_LAUNCH();
  std::swap(ps->pos, ps->pos2);
  std::swap(ps->vel[ps->cur], ps->vel2);
  std::swap(ps->mass, ps->mass2);
  std::swap(ps->acc, ps->acc2);
  std::swap(ps->orig, ps->orig2);
  double dense = 1.0;
  for (int d = 0; d < 3; ++d) dense *= (double)cs.ext[d];
  TSPH_REQUIRE(dense <= (double)kMaxDenseCells, TSPH_ERR_ARG, "periodic grid of %.3g cells exceeds the dense "
               "range table (%lld cells)", dense, (long long)kMaxDenseCells);
  ps->ncells_dense = (int64_t)dense;
  for (int d = 0; d < 3; ++d) { ps->cell_ext[d] = cs.ext[d]; ps->cell_stride[d] = cs.stride[d]; }
  TSPH_TRY(ps->cell_R.ensure(sizeof(int4) * (size_t)ps->ncells_dense));
  TSPH_TRY(ps->entries.ensure(sizeof(unsigned long long)));
  k_cell_table_clear<<<grid_for(ps->ncells_dense), kBlock, 0, stream>>>(ps->cell_R.as<int4>(), ps->ncells_dense);
  k_cell_table<3><<<g, kBlock, 0, stream>>>(ps->cells2.as<int64_t>(), n, n, cs.stride[0], cs.stride[1],
                                            cs.stride[2], ps->cell_R.as<int4>());
  TSPH_CHECK_LAUNCH();
  // per-row lists at the build radius (minimum image) in fixed slots; their
  // entry count is the periodic_pairs entry count
  TSPH_TRY((build_slot_lists<3, true>(psim_cells(ps), ps->pos.as<double>(), n, n, rc * rc, ps->prm.box,
                                      ps->slot_cnt, ps->slot_idx, ps->slot_max, ps->slot_K,
                                      ps->entries.as<unsigned long long>(), stream, &ps->slot_posf,
                                      fmax(fmax(ps->prm.box[0], ps->prm.box[1]), ps->prm.box[2]))));
  TSPH_TRY(ps->sup_idx.ensure(sizeof(int32_t) * (size_t)((n + 63) / 64) * 64 * (size_t)ps->slot_K));
  TSPH_TRY(ps->sup_cnt.ensure(sizeof(int32_t) * (size_t)n));
  TSPH_CHECK_HIP(hipMemcpyAsync(ps->anchor.ptr, ps->pos.ptr, 24 * n, hipMemcpyDeviceToDevice, stream));
  // fresh lists: the sweeps run again, the displacement counts from here
  TSPH_CHECK_HIP(hipMemsetAsync(ps->stale.ptr, 0, sizeof(unsigned int), stream));
  TSPH_CHECK_HIP(hipMemsetAsync(ps->disp.ptr, 0, sizeof(unsigned long long), stream));
  ps->rebuilds++;
  ps->built = true;
  return TSPH_OK;
}

template <typename S>
int psim_eval(tsph_psim* ps, bool kick, hipStream_t stream) {
  ThreedArgs<S> a = make_threed_args<S>(ps->q, ps->n, ps->n);
  a.pos = ps->pos.as<double>();
  a.vel = ps->vel[ps->cur].as<double>();
  a.mass = ps->mass.as<S>();
  a.mass_lo = sizeof(S) == 4 ? ps->mass.as<S>() + ps->n : nullptr;
  a.skip = ps->stale.as<unsigned int>();
  const NbSlots nb = psim_slots(ps);
  NbSlots sup = nb;
  sup.idx = ps->sup_idx.as<int32_t>();
  sup.cnt = ps->sup_cnt.as<int32_t>();
  if (kick) {
    TSPH_TRY((threed_sweeps<S, true>(a, ps->q, psim_scratch<S>(ps), nb, ps->orig.as<int64_t>(),
                                     ps->acc.as<S>(), 0.5 * ps->prm.dt, ps->vel[ps->cur ^ 1].as<double>(), stream,
                                     ps->uniform_mass, &sup)));
    k3_carry_vel<<<grid_for(3 * ps->n), kBlock, 0, stream>>>(ps->stale.as<unsigned int>(),
                                                             ps->vel[ps->cur].as<double>(),
                                                             ps->vel[ps->cur ^ 1].as<double>(), 3 * ps->n);
    TSPH_CHECK_LAUNCH();
    ps->cur ^= 1;
  } else {
    TSPH_TRY((threed_sweeps<S, false>(a, ps->q, psim_scratch<S>(ps), nb, ps->orig.as<int64_t>(),
                                      ps->acc.as<S>(), 0.0, nullptr, stream, ps->uniform_mass, &sup)));
  }
  return TSPH_OK;
}

// Leapfrog steps (threed/solver.py:285-299) without a host wait per step:
// each step is queued as kick-drift, the device skin test, the evaluation
// and the second half-kick, and an event; the host reads a step's skin flag
// two steps later.  A step whose drift went past half the skin has its
// evaluation skipped on the device, and so has every step queued after it;
// the host then rebuilds the lists from that step's drifted positions, runs
// its evaluation, and queues the later steps again.  The rebuild schedule is
// the sequential one's: the lists are rebuilt exactly when a drift goes past
// half the skin, before that step's evaluation.
template <typename S>
int psim_run(tsph_psim* ps, int64_t nsteps, hipStream_t stream) {
  if (!ps->built) TSPH_TRY(psim_rebuild<S>(ps, stream));
  if (!ps->have_acc) { TSPH_TRY(psim_eval<S>(ps, false, stream)); ps->have_acc = true; }
  const double* b = ps->prm.box;
  constexpr int R = tsph_psim::kRingP;
  const int depth = 2;
  int64_t next = 0, done = 0;
  while (done < nsteps) {
    while (next < nsteps && next - done < depth) {
      k3_kick_drift<S><<<grid_for(ps->n), kBlock, 0, stream>>>(ps->pos.as<double>(), ps->vel[ps->cur].as<double>(),
          ps->acc.as<S>(), ps->n, ps->prm.dt, b[0], b[1], b[2], ps->anchor.as<double>(),
          ps->disp.as<unsigned long long>(), ps->stale.as<unsigned int>());
      k3_skin_test<<<1, 64, 0, stream>>>(ps->disp.as<unsigned long long>(), ps->skin, ps->stale.as<unsigned int>(),
                                         ps->ring + next % R);
      TSPH_CHECK_LAUNCH();
      TSPH_TRY(psim_eval<S>(ps, true, stream));
      TSPH_CHECK_HIP(hipEventRecord(ps->ev[next % R], stream));
      ++next;
    }
    TSPH_CHECK_HIP(hipEventSynchronize(ps->ev[done % R]));
    if (ps->ring[done % R]) {
      // the steps queued after it did nothing: let them finish, rebuild,
      // evaluate this step, and queue the rest again
      for (int64_t k = done + 1; k < next; ++k) TSPH_CHECK_HIP(hipEventSynchronize(ps->ev[k % R]));
      TSPH_TRY(psim_rebuild<S>(ps, stream));
      TSPH_TRY(psim_eval<S>(ps, true, stream));
      next = done + 1;
    }
    ps->steps++;
    ++done;
  }
  return TSPH_OK;
}

}  // namespace tsph

// ---------------------------------------------------------------- allocation-free conservation row
// Per-block partials of tsph_conservation_row_noalloc live in this file-scope
// array: DG_N doubles per block of kBlock rows plus the DG_N of the reduced
// row, 8 MiB in all (up to kNoallocMaxRows rows).  Host code reaches it only
// through the device address hipGetSymbolAddress returns, once per device.
constexpr int64_t kNoallocDoubles = 1 << 20;
constexpr int64_t kNoallocMaxBlocks = kNoallocDoubles / DG_N - 1;
constexpr int64_t kNoallocMaxRows = kNoallocMaxBlocks * kBlock;
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once
int noalloc_scratch(double** out) {
  int dev = -1;
  TSPH_CHECK_HIP(hipGetDevice(&dev));
  TSPH_REQUIRE(dev >= 0 && dev < kMaxDevices, TSPH_ERR_ARG, "device %d outside the %d this library tracks", dev,
               kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TSPH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TSPH_REQUIRE(p != nullptr, TSPH_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return TSPH_OK;
}
}  // namespace

// ====================================================================== C ABI
extern "C" {

int tsph_threed_accelerations(tsph_ctx* ctx, const tsph_threed_params* prm, const double* pos,
                              const double* vel, const void* mass, void* acc, void* rho, void* p,
                              hipStream_t stream) {
  TSPH_REQUIRE(ctx && prm, TSPH_ERR_ARG, "ctx and params must not be NULL");
  TSPH_REQUIRE(prm->n >= 0 && prm->n_fluid >= 0 && prm->n_fluid <= prm->n, TSPH_ERR_ARG,
               "bad particle counts n=%lld n_fluid=%lld", (long long)prm->n, (long long)prm->n_fluid);
  TSPH_REQUIRE(prm->precision == TSPH_F64 || prm->precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(ctx->nb_targets == prm->n && ctx->wnb_targets == 0, TSPH_ERR_ARG,
               "neighbour lists hold %lld targets; accelerations needs lists for all %lld particles "
               "(tsph_nlist_from_pairs with all_targets = 1)", (long long)ctx->nb_targets, (long long)prm->n);
  TSPH_REQUIRE(prm->h > 0.0 && prm->rho0 > 0.0 && prm->gamma > 0.0, TSPH_ERR_ARG,
               "h, rho0 and gamma must be positive");
  TSPH_REQUIRE(prm->wall_model == 0 || prm->wall_model == 1, TSPH_ERR_ARG, "wall_model must be 0 or 1");
  TSPH_REQUIRE(prm->viscous_order == 0 || prm->viscous_order == 2 || prm->viscous_order == 4, TSPH_ERR_ARG,
               "viscous_order must be 2 or 4");
  TSPH_REQUIRE(prm->diffusion_scale <= 0.0 || prm->dt > 0.0, TSPH_ERR_ARG,
               "density_diffusion > 0 requires a positive dt");
  if (prm->periodic)
    for (int d = 0; d < 3; ++d) TSPH_REQUIRE(prm->box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive");
  if (prm->n > 0) TSPH_REQUIRE(pos && vel && mass && acc, TSPH_ERR_ARG, "NULL array");
  return prm->precision == TSPH_F64 ? threed_accel<double>(ctx, prm, pos, vel, mass, acc, rho, p, stream)
                                    : threed_accel<float>(ctx, prm, pos, vel, mass, acc, rho, p, stream);
}

int tsph_threed_operator(tsph_ctx* ctx, const tsph_threed_params* prm, int op, const double* pos,
                         const double* vel, const void* mass, const void* rho, const void* p, void* out,
                         hipStream_t stream) {
  TSPH_REQUIRE(ctx && prm, TSPH_ERR_ARG, "ctx and params must not be NULL");
  TSPH_REQUIRE(op >= 0 && op <= 4, TSPH_ERR_ARG, "unknown operator %d", op);
  TSPH_REQUIRE(prm->n >= 0 && prm->n_fluid >= 0 && prm->n_fluid <= prm->n, TSPH_ERR_ARG,
               "bad particle counts n=%lld n_fluid=%lld", (long long)prm->n, (long long)prm->n_fluid);
  TSPH_REQUIRE(prm->precision == TSPH_F64 || prm->precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(ctx->nb_targets == prm->n && ctx->wnb_targets == 0, TSPH_ERR_ARG,
               "neighbour lists hold %lld targets; the operator needs lists for all %lld particles "
               "(tsph_nlist_from_pairs with all_targets = 1)", (long long)ctx->nb_targets, (long long)prm->n);
  TSPH_REQUIRE(prm->h > 0.0, TSPH_ERR_ARG, "h must be positive");
  TSPH_REQUIRE(prm->viscous_order == 0 || prm->viscous_order == 2 || prm->viscous_order == 4, TSPH_ERR_ARG,
               "viscous_order must be 2 or 4");
  if (prm->periodic)
    for (int d = 0; d < 3; ++d) TSPH_REQUIRE(prm->box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive");
  if (prm->n > 0) {
    TSPH_REQUIRE(pos && mass && out, TSPH_ERR_ARG, "NULL array");
    TSPH_REQUIRE(op != 2 || vel, TSPH_ERR_ARG, "viscous_acceleration needs vel");
    TSPH_REQUIRE(op == 0 || rho, TSPH_ERR_ARG, "this operator needs rho");
    TSPH_REQUIRE((op != 1 && op != 3) || p, TSPH_ERR_ARG, "this operator needs the pressure");
  }
  return prm->precision == TSPH_F64
             ? threed_operator<double>(ctx, prm, op, pos, vel, mass, rho, p, out, stream)
             : threed_operator<float>(ctx, prm, op, pos, vel, mass, rho, p, out, stream);
}

int tsph_conservation_row(tsph_ctx* ctx, const double* pos, const double* vel, const void* mass,
                          const void* rho, int64_t n, int precision, double rho0, double sound_speed, double gamma,
                          const double* box, double wavenumber, double* out, hipStream_t stream) {
  TSPH_REQUIRE(ctx && out && n >= 0, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(precision == TSPH_F64 || precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(n == 0 || (pos && vel && mass && rho), TSPH_ERR_ARG, "NULL array");
  TSPH_REQUIRE(rho0 > 0.0 && gamma > 0.0 && gamma != 1.0, TSPH_ERR_ARG, "rho0 > 0 and gamma > 0, != 1 required");
  double b[3] = {1.0, 1.0, 1.0};
  if (box) for (int d = 0; d < 3; ++d) { TSPH_REQUIRE(box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive"); b[d] = box[d]; }
  const unsigned g = n > 0 ? grid_for(n) : 1;
  TSPH_TRY(ctx->s_red.ensure(sizeof(double) * DG_N * (g + 1)));
  double* red = ctx->s_red.as<double>();
  if (precision == TSPH_F64)
    k3_diag<double><<<g, kBlock, 0, stream>>>(pos, vel, (const double*)mass, (const double*)rho, n, rho0,
                                              sound_speed, gamma, b[0], b[1], b[2], wavenumber, red);
  else
    k3_diag<float><<<g, kBlock, 0, stream>>>(pos, vel, (const float*)mass, (const float*)rho, n, rho0,
                                             sound_speed, gamma, b[0], b[1], b[2], wavenumber, red);
  k3_diag_reduce<<<1, 256, 0, stream>>>(red, g, red + (size_t)DG_N * g);
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipMemcpyAsync(out, red + (size_t)DG_N * g, sizeof(double) * DG_N, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  return TSPH_OK;
}

int tsph_conservation_row_noalloc(const double* pos, const double* vel, const void* mass, const void* rho,
                                  int64_t n, int precision, double rho0, double sound_speed, double gamma,
                                  const double* box, double wavenumber, double* out, hipStream_t stream) {
  TSPH_REQUIRE(n >= 0 && n <= kNoallocMaxRows, TSPH_ERR_ARG,
               "%lld rows exceed the %lld the allocation-free conservation row holds", (long long)n,
               (long long)kNoallocMaxRows);
  TSPH_REQUIRE(out, TSPH_ERR_ARG, "out must not be NULL");
  TSPH_REQUIRE(precision == TSPH_F64 || precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(n == 0 || (pos && vel && mass && rho), TSPH_ERR_ARG, "NULL array");
  TSPH_REQUIRE(rho0 > 0.0 && gamma > 0.0 && gamma != 1.0, TSPH_ERR_ARG, "rho0 > 0 and gamma > 0, != 1 required");
  double b[3] = {1.0, 1.0, 1.0};
  if (box) for (int d = 0; d < 3; ++d) { TSPH_REQUIRE(box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive"); b[d] = box[d]; }
  double* red = nullptr;
  TSPH_TRY(noalloc_scratch(&red));
  // the same kernels and grid as tsph_conservation_row: every block writes its
  // partials before the reduction reads them, so nothing needs clearing
  const unsigned g = n > 0 ? grid_for(n) : 1;
  if (precision == TSPH_F64)
    k3_diag<double><<<g, kBlock, 0, stream>>>(pos, vel, (const double*)mass, (const double*)rho, n, rho0,
                                              sound_speed, gamma, b[0], b[1], b[2], wavenumber, red);
  else
    k3_diag<float><<<g, kBlock, 0, stream>>>(pos, vel, (const float*)mass, (const float*)rho, n, rho0,
                                             sound_speed, gamma, b[0], b[1], b[2], wavenumber, red);
  k3_diag_reduce<<<1, 256, 0, stream>>>(red, g, red + (size_t)DG_N * g);
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipMemcpyAsync(out, red + (size_t)DG_N * g, sizeof(double) * DG_N, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  return TSPH_OK;
}

int tsph_psim_destroy(tsph_psim* ps) {
  if (!ps) return TSPH_OK;
  DevBuf* b[] = {&ps->pos, &ps->vel[0], &ps->vel[1], &ps->mass, &ps->acc, &ps->orig, &ps->anchor, &ps->disp,
                 &ps->pos2, &ps->vel2, &ps->mass2, &ps->acc2, &ps->orig2, &ps->cells2, &ps->part, &ps->diag,
                 &ps->rho_raw, &ps->rho, &ps->rho_u, &ps->p, &ps->moment, &ps->gr, &ps->rho_lo, &ps->pack,
                 &ps->cell_R, &ps->entries, &ps->slot_idx, &ps->slot_cnt, &ps->slot_max, &ps->slot_posf,
                 &ps->sup_idx, &ps->sup_cnt, &ps->fourth, &ps->lap1};
  for (DevBuf* x : b) x->release();
  ps->stale.release();
  for (int r = 0; r < tsph_psim::kRingP; ++r)
    if (ps->ev[r]) (void)hipEventDestroy(ps->ev[r]);
  if (ps->ring) (void)hipHostFree(ps->ring);
  if (ps->disp_host) (void)hipHostFree(ps->disp_host);
  if (ps->ctx) tsph_ctx_destroy(ps->ctx);
  delete ps;
  return TSPH_OK;
}

int tsph_psim_create(const tsph_threed_params* prm, double skin, const double* pos, const double* vel,
                     const void* mass, hipStream_t stream, tsph_psim** out) {
  TSPH_REQUIRE(prm && out && pos && vel && mass, TSPH_ERR_ARG, "NULL argument");
  *out = nullptr;
  TSPH_REQUIRE(prm->periodic && prm->n > 0 && prm->n_fluid == prm->n, TSPH_ERR_ARG,
               "the periodic stepper advances a triply periodic fluid without walls");
  TSPH_REQUIRE(prm->n < (1ll << 31) - 1, TSPH_ERR_ARG, "too many particles for int32 neighbour lists");
  TSPH_REQUIRE(prm->precision == TSPH_F64 || prm->precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(skin > 0.0 && prm->dt > 0.0 && prm->h > 0.0, TSPH_ERR_ARG, "skin, dt and h must be positive");
  TSPH_REQUIRE(prm->viscous_order == 0 || prm->viscous_order == 2 || prm->viscous_order == 4, TSPH_ERR_ARG,
               "viscous_order must be 2 or 4");
  for (int d = 0; d < 3; ++d)
    TSPH_REQUIRE(prm->box[d] > 0.0, TSPH_ERR_ARG, "box edges must be positive");
  tsph_psim* ps = new tsph_psim();
  ps->prm = *prm;
  ps->q = threed_prm(*prm);
  ps->skin = skin;
  ps->n = prm->n;
  ps->s = prm->precision == TSPH_F64 ? 8 : 4;
  int rc = tsph_ctx_create(&ps->ctx);
  const size_t n = (size_t)ps->n;
  const size_t s = ps->s;
  DevBuf* bufs[] = {&ps->pos, &ps->vel[0], &ps->vel[1], &ps->vel2, &ps->pos2, &ps->anchor, &ps->cells2,
                    &ps->mass, &ps->mass2, &ps->acc, &ps->acc2, &ps->orig, &ps->orig2, &ps->disp,
                    &ps->part, &ps->diag, &ps->rho_raw, &ps->rho, &ps->rho_u, &ps->p, &ps->moment, &ps->gr,
                    &ps->rho_lo, &ps->pack};
  const size_t sizes[] = {24 * n, 24 * n, 24 * n, 24 * n, 24 * n, 24 * n, 24 * n,
                          (s == 4 ? 2 : 1) * s * (n + 1), (s == 4 ? 2 : 1) * s * (n + 1),
                          3 * s * (n + 1), 3 * s * (n + 1), 8 * n, 8 * n, 64,
                          8 * DG_N * ((size_t)grid_for(ps->n) + 1), 8 * DG_N, s * (n + 1), s * (n + 1), s * (n + 1),
                          s * (n + 1), s * (n + 1), 3 * s * (n + 1), s * (n + 1), 4 * s * (n + 1)};
  for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]) && rc == TSPH_OK; ++k) rc = bufs[k]->ensure(sizes[k]);
  if (rc == TSPH_OK && ps->q.order4) rc = ps->fourth.ensure(s * (n + 1));
  if (rc == TSPH_OK && ps->q.order4) rc = ps->lap1.ensure(3 * s * (n + 1));
  if (rc == TSPH_OK && hipHostMalloc((void**)&ps->disp_host, 8, 0) != hipSuccess) {
    ps->disp_host = nullptr;
    set_error("hipHostMalloc of the displacement word failed");
    rc = TSPH_ERR_HIP;
  }
  if (rc == TSPH_OK) rc = ps->stale.ensure(sizeof(unsigned int));
  if (rc == TSPH_OK && hipHostMalloc((void**)&ps->ring, sizeof(unsigned int) * tsph_psim::kRingP, 0) != hipSuccess) {
    ps->ring = nullptr;
    set_error("hipHostMalloc of the skin-test ring failed");
    rc = TSPH_ERR_HIP;
  }
  for (int r = 0; r < tsph_psim::kRingP && rc == TSPH_OK; ++r)
    if (hipEventCreateWithFlags(&ps->ev[r], hipEventDisableTiming) != hipSuccess) {
      ps->ev[r] = nullptr;
      set_error("hipEventCreate failed");
      rc = TSPH_ERR_HIP;
    }
  if (rc == TSPH_OK && hipMemsetAsync(ps->stale.ptr, 0, sizeof(unsigned int), stream) != hipSuccess) {
    set_error("hipMemsetAsync failed");
    rc = TSPH_ERR_HIP;
  }
  if (rc != TSPH_OK) { tsph_psim_destroy(ps); return rc; }
  hipError_t e = hipMemcpyAsync(ps->pos.ptr, pos, 24 * n, hipMemcpyDeviceToDevice, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ps->vel[0].ptr, vel, 24 * n, hipMemcpyDeviceToDevice, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ps->mass.ptr, mass, s * n, hipMemcpyDeviceToDevice, stream);
  // FP32 runs: the masses' low parts (the second half of the buffer) start at
  // zero; tsph_psim_set_masses_f64 fills them
  if (e == hipSuccess && s == 4) e = hipMemsetAsync(ps->mass.as<float>() + n, 0, s * n, stream);
  if (e == hipSuccess) e = hipMemsetAsync(ps->acc.ptr, 0, 3 * s * n, stream);
  if (e != hipSuccess) {
    set_error("periodic stepper init: %s", hipGetErrorString(e));
    tsph_psim_destroy(ps);
    return TSPH_ERR_HIP;
  }
  k3_iota<<<grid_for(ps->n), kBlock, 0, stream>>>(ps->orig.as<int64_t>(), ps->n);
  // disp (64 bytes) is free until the first step: it takes the equal-mass flag
  if (e == hipSuccess) e = hipMemsetAsync(ps->disp.ptr, 0, sizeof(unsigned int), stream);
  if (s == 8)
    k3_mass_differs<double><<<grid_for(ps->n), kBlock, 0, stream>>>(ps->mass.as<double>(), ps->n, ps->disp.as<unsigned int>());
  else
    k3_mass_differs<float><<<grid_for(ps->n), kBlock, 0, stream>>>(ps->mass.as<float>(), ps->n, ps->disp.as<unsigned int>());
  unsigned int differs = 1;
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&differs, ps->disp.ptr, sizeof(differs), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  ps->uniform_mass = differs == 0;
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("periodic stepper init: %s", hipGetErrorString(e));
    tsph_psim_destroy(ps);
    return TSPH_ERR_HIP;
  }
  *out = ps;
  return TSPH_OK;
}

int tsph_psim_set_masses_f64(tsph_psim* ps, const double* mass, hipStream_t stream) {
  TSPH_REQUIRE(ps && mass, TSPH_ERR_ARG, "NULL argument");
  const int64_t n = ps->n;
  if (ps->s == 8)
    k3_set_masses<double><<<grid_for(n), kBlock, 0, stream>>>(mass, ps->orig.as<int64_t>(), n,
                                                              ps->mass.as<double>(), 0);
  else
    k3_set_masses<float><<<grid_for(n), kBlock, 0, stream>>>(mass, ps->orig.as<int64_t>(), n,
                                                             ps->mass.as<float>(), n);
  TSPH_CHECK_LAUNCH();
  // the equal-mass flag of the sweeps, from the new masses (diag is scratch
  // until the next diagnostics call writes it)
  unsigned int* flag = ps->diag.as<unsigned int>();
  TSPH_CHECK_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), stream));
  k3_mass_differs<double><<<grid_for(n), kBlock, 0, stream>>>(mass, n, flag);
  TSPH_CHECK_LAUNCH();
  unsigned int differs = 1;
  TSPH_CHECK_HIP(hipMemcpyAsync(&differs, flag, sizeof(differs), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  ps->uniform_mass = differs == 0;
  return TSPH_OK;
}

int tsph_psim_step(tsph_psim* ps, int64_t nsteps, hipStream_t stream) {
  TSPH_REQUIRE(ps && nsteps >= 0, TSPH_ERR_ARG, "bad arguments");
  return ps->s == 8 ? psim_run<double>(ps, nsteps, stream) : psim_run<float>(ps, nsteps, stream);
}

int tsph_psim_diagnostics(tsph_psim* ps, double wavenumber, double* out, hipStream_t stream) {
  TSPH_REQUIRE(ps && out, TSPH_ERR_ARG, "bad arguments");
  TSPH_TRY(ps->s == 8 ? psim_run<double>(ps, 0, stream) : psim_run<float>(ps, 0, stream));
  const unsigned g = grid_for(ps->n);
  const double* b = ps->prm.box;
  const tsph_threed_params& q = ps->prm;
  if (ps->s == 8)
    k3_diag<double><<<g, kBlock, 0, stream>>>(ps->pos.as<double>(), ps->vel[ps->cur].as<double>(),
        ps->mass.as<double>(), ps->rho.as<double>(), ps->n, q.rho0, q.sound_speed, q.gamma, b[0], b[1], b[2],
        wavenumber, ps->part.as<double>());
  else
    k3_diag<float><<<g, kBlock, 0, stream>>>(ps->pos.as<double>(), ps->vel[ps->cur].as<double>(),
        ps->mass.as<float>(), ps->rho.as<float>(), ps->n, q.rho0, q.sound_speed, q.gamma, b[0], b[1], b[2],
        wavenumber, ps->part.as<double>());
  k3_diag_reduce<<<1, 256, 0, stream>>>(ps->part.as<double>(), g, ps->diag.as<double>());
  TSPH_CHECK_LAUNCH();
  TSPH_CHECK_HIP(hipMemcpyAsync(out, ps->diag.ptr, 8 * DG_N, hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  return TSPH_OK;
}

int tsph_psim_read(tsph_psim* ps, double* pos, double* vel, void* rho, void* acc, hipStream_t stream) {
  TSPH_REQUIRE(ps, TSPH_ERR_ARG, "NULL stepper");
  TSPH_TRY(ps->s == 8 ? psim_run<double>(ps, 0, stream) : psim_run<float>(ps, 0, stream));
  const unsigned g = grid_for(ps->n);
  if (ps->s == 8)
    k3_unpermute<double><<<g, kBlock, 0, stream>>>(ps->orig.as<int64_t>(), ps->n, ps->pos.as<double>(),
        ps->vel[ps->cur].as<double>(), ps->rho.as<double>(), ps->acc.as<double>(), pos, vel, (double*)rho,
        (double*)acc);
  else
    k3_unpermute<float><<<g, kBlock, 0, stream>>>(ps->orig.as<int64_t>(), ps->n, ps->pos.as<double>(),
        ps->vel[ps->cur].as<double>(), ps->rho.as<float>(), ps->acc.as<float>(), pos, vel, (float*)rho,
        (float*)acc);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

int tsph_psim_stats(tsph_psim* ps, int64_t* counters) {
  TSPH_REQUIRE(ps && counters, TSPH_ERR_ARG, "bad arguments");
  counters[0] = ps->steps;
  counters[1] = ps->rebuilds;
  unsigned long long e = 0;
  if (ps->built) TSPH_CHECK_HIP(hipMemcpy(&e, ps->entries.ptr, sizeof(e), hipMemcpyDeviceToHost));
  counters[2] = (int64_t)e;
  return TSPH_OK;
}

}  // extern "C"
