// Weakly-compressible SPH right-hand side with the Morris viscous Laplacian:
// droplet_impact.rhs (droplet_impact.py:810-1228, D = 2) and
// dam_break_3d.rhs_3d (dam_break_3d.py:290-422, D = 3).
//
// The reference scatters every unordered pair to both ends with index_add.
// Here every target particle gathers over its own neighbour list instead:
// the pair term a particle receives is the same expression, evaluated from
// its side, so no two threads ever write the same address and the result is
// independent of scheduling.  Sweeps:
//   A  density (+ self term, floor at 0.1 rho0), Tait pressure     [fluid]
//   A' density diffusion (Molteni, or Antuono with a renormalised
//      gradient pre-pass), then Tait                                 [fluid]
//   W  Adami wall pressure / Tait-consistent wall density            [wall]
//   B  pressure force + Morris viscosity + gravity                   [fluid]
//   R  wall reactions (2-D split loads / rhs_3d wall force)          [wall]
#include <hip/hip_runtime.h>

#include "tsph_common.hpp"
#include "tsph_internal.hpp"
#include "tsph_wcsph.hpp"

namespace tsph {

// ------------------------------------------------------------------ sweep A
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_density(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ idx, int write_eos) {
  const int64_t nb = gridDim.x;
  const int64_t i = xcd_block(blockIdx.x, nb) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> acc, kt;
  const int64_t e0 = off[i], e1 = off[i + 1];
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t j = idx[e];
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double df = sep(a, xi[d] - a.pos[j * D + d], d);
      r2 += df * df;
    }
    if (r2 > a.skip2d) continue;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    acc.add((double)a.mass[j] * w);
    if (a.shepard) kt.add(w);
  }
  const double floor_d = (double)a.rho_floor;
  double rho_d = hi(acc) + (double)a.mass[i] * a.kd.w0d;
  rho_d = rho_d > floor_d ? rho_d : floor_d;
  if (a.dens_off) {                         // droplet_impact.py:936-943
    rho_d = rho_d + (double)a.dens_off[i];
    rho_d = rho_d > floor_d ? rho_d : floor_d;
  }
  if (a.shepard) {                          // Shepard renormalisation, :954-964
    double t = hi(kt) + a.kd.w0d;
    t = t > 1.0e-12 ? t : 1.0e-12;
    rho_d = rho_d * (a.lattice_sum / t);
    rho_d = rho_d > floor_d ? rho_d : floor_d;
  }
  const S rho = store_rho<S>(rho_d, a.rho, a.rho_lo, i);
  if (write_eos) {
    const S p = tait_hi<S>(rho_d, a.rho0d, a.bulk, a.gamma, a.ratio_floor);
    a.p[i] = p;
    a.pr2[i] = p / (rho * rho);
    a.vol[i] = a.mass[i] / rho;
  }
}

// ------------------------------------------------------------------ sweep A'
// Renormalised density gradient (density_diffusion.py:95-142) over fluid-fluid
// pairs: G = sum V_j (rho_j - rho_i) grad_i W, M = -sum V_j delta (x) grad_i W,
// <grad rho> = f(M) G with f(l) = 1/max(l, floor * trace/D).
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_renorm_gradient(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ idx,
                                                            S* __restrict__ grad_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  const S rhoi = a.rho[i];
  Sum<S> G[D], M[D * D];
  for (int64_t e = off[i]; e < off[i + 1]; ++e) {
    const int64_t j = idx[e];
    if (j >= a.n) continue;
    S df[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)sep(a, xi[d] - a.pos[j * D + d], d); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) continue;
    S w, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
    const S vj = a.mass[j] / a.rho[j];
    const S diff = a.rho[j] - rhoi;
#pragma unroll
    for (int r = 0; r < D; ++r) {
      const S gr = df[r] * g;
      G[r].add((vj * diff) * gr);
#pragma unroll
      for (int c = 0; c < D; ++c) M[r * D + c].add(vj * (-(df[r] * (df[c] * g))));
    }
  }
  double m[9], gv[3], out[3];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    gv[r] = (double)G[r].get();
#pragma unroll
    for (int c = 0; c < D; ++c) m[r * D + c] = (double)M[r * D + c].get();
  }
  double tr = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) tr += m[d * D + d];
  if (tr <= 0.0) {
#pragma unroll
    for (int d = 0; d < D; ++d) grad_out[i * D + d] = S(0);
    return;
  }
  regularised_solve<D>(m, gv, (double)a.eig_floor, out);
#pragma unroll
  for (int d = 0; d < D; ++d) grad_out[i * D + d] = (S)out[d];
}

template <int D, typename S, int FORM>
__global__ void __launch_bounds__(kBlock) k_diffusion(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ idx,
                                                      const S* __restrict__ rho_in,
                                                      const S* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
  for (int64_t e = off[i]; e < off[i + 1]; ++e) {
    const int64_t j = idx[e];
    if (j >= a.n) continue;
    S df[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)sep(a, xi[d] - a.pos[j * D + d], d); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) continue;
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
    const S coupling = contrast * radial / (r2 + S(1e-12));
    rate.add(-a.diff_scale * (a.mass[j] / rho_in[j]) * coupling);
  }
  double rho_d = (double)rhoi + (a.rho_lo ? (double)a.rho_lo[i] : 0.0) + (double)a.dt * hi(rate);
  rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  const S rho = (S)rho_d;
  const S p = tait_hi<S>(rho_d, a.rho0d, a.bulk, a.gamma, a.ratio_floor);
  a.rho[i] = rho;
  a.p[i] = p;
  a.pr2[i] = p / (rho * rho);
  a.vol[i] = a.mass[i] / rho;
}

// ------------------------------------------------------------------ sweep W
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_wall_pressure(WcsphArgs<S> a, const int64_t* __restrict__ woff,
                                                          const int32_t* __restrict__ widx,
                                                          S* __restrict__ wall_p_out,
                                                          S* __restrict__ wall_rho_out) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nw) return;
  const int64_t iw = a.n + w;
  double xw[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xw[d] = a.pos[iw * D + d];
  double aw[D];
#pragma unroll
  for (int d = 0; d < D; ++d) aw[d] = a.wall_accel ? a.wall_accel[w * D + d] : 0.0;
  Sum<S> num, den;
  for (int64_t e = woff[w]; e < woff[w + 1]; ++e) {
    const int64_t f = widx[e];
    S rw[D], r2 = S(0);                       // r_w - r_f
#pragma unroll
    for (int d = 0; d < D; ++d) { rw[d] = (S)sep(a, xw[d] - a.pos[f * D + d], d); r2 += rw[d] * rw[d]; }
    if (r2 > a.skip2) continue;
    S wgt, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, wgt, g);
    S body = -a.gravity * rw[1];
    if (a.wall_accel) body = body - ((S)aw[0] * rw[0] + (S)aw[1] * rw[1]);
    S src = a.p[f];
    if (a.gravity != S(0)) src = src + a.rho[f] * body;
    num.add(src * wgt);
    den.add(wgt);
  }
  const S dd = den.get() > S(1e-30) ? den.get() : S(1e-30);
  S pw = num.get() / dd;
  pw = pw > S(0) ? pw : S(0);
  S rw = a.rho0 * pow_s<S>(pw / a.bulk + S(1), a.inv_gamma);
  rw = rw > a.rho0 ? rw : a.rho0;
  a.p[iw] = pw;
  a.rho[iw] = rw;
  a.pr2[iw] = pw / (rw * rw);
  a.vol[iw] = a.mass[iw] / rw;
  if (wall_p_out) wall_p_out[w] = pw;
  if (wall_rho_out) wall_rho_out[w] = rw;
}

// ------------------------------------------------------------------ sweep B
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_force(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ idx,
                                                  S* __restrict__ acc_out) {
  const int64_t nb = gridDim.x;
  const int64_t i = xcd_block(blockIdx.x, nb) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D], vi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { xi[d] = a.pos[i * D + d]; vi[d] = a.vel[i * D + d]; }
  const S pri = a.pr2[i];
  const S nui = a.nu_field ? a.nu_field[i] : S(0);
  Sum<S> ap[D], av[D];
  for (int64_t e = off[i]; e < off[i + 1]; ++e) {
    const int64_t j = idx[e];
    S df[D], r2 = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)sep(a, xi[d] - a.pos[j * D + d], d); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) continue;
    S w, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
    const S mj = a.mass[j];
    const S term = pri + a.pr2[j];
    const S cp = -mj * term;
#pragma unroll
    for (int d = 0; d < D; ++d) ap[d].add(cp * (df[d] * g));
    // Morris viscosity: fluid-fluid always; fluid-wall with the wall_friction blend
    const bool fluid_j = j < a.n;
    const S blend = fluid_j ? S(1) : a.friction;
    if (blend != S(0)) {
      S dg = S(0);
#pragma unroll
      for (int d = 0; d < D; ++d) dg += df[d] * (df[d] * g);
      const S relu = -dg > S(0) ? -dg : S(0);
      const S lw = relu / (r2 + a.eta2);
      // a viscosity field: mean of the two members on fluid pairs, the fluid
      // member's on fluid-wall pairs (droplet_impact.py:1168-1173)
      const S nu2 = a.nu_field ? S(2) * (fluid_j ? S(0.5) * (nui + a.nu_field[j]) : nui) : a.nu2;
      const S cv = -(nu2 * blend) * (a.vol[j] * lw);
#pragma unroll
      for (int d = 0; d < D; ++d) av[d].add(cv * (S)(vi[d] - a.vel[j * D + d]));
    }
  }
  S acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = (ap[d].get() + av[d].get()) + (d == 1 ? -a.gravity : S(0));
  if (acc_out) {
#pragma unroll
    for (int d = 0; d < D; ++d) acc_out[i * D + d] = acc[d];
  }
}

// ------------------------------------------------------------------ sweep R
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_wall_loads(WcsphArgs<S> a, const int64_t* __restrict__ woff,
                                                       const int32_t* __restrict__ widx,
                                                       S* __restrict__ reaction) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nw) return;
  const int64_t iw = a.n + w;
  double xw[D], vw[D];
#pragma unroll
  for (int d = 0; d < D; ++d) { xw[d] = a.pos[iw * D + d]; vw[d] = a.vel[iw * D + d]; }
  const S prw = a.pr2[iw], mw = a.mass[iw], volw = a.vol[iw];
  Sum<S> r[D];
  for (int64_t e = woff[w]; e < woff[w + 1]; ++e) {
    const int64_t f = widx[e];
    S df[D], r2 = S(0);                      // delta = x_f - x_w
#pragma unroll
    for (int d = 0; d < D; ++d) { df[d] = (S)sep(a, a.pos[f * D + d] - xw[d], d); r2 += df[d] * df[d]; }
    if (r2 > a.skip2) continue;
    S wgt, g;
    cubic_wg<S>(sqrt_s<S>(r2), a.k, wgt, g);
    const S mf = a.mass[f];
    const S term = a.pr2[f] + prw;
    S dg = S(0);
#pragma unroll
    for (int d = 0; d < D; ++d) dg += df[d] * (df[d] * g);
    const S relu = -dg > S(0) ? -dg : S(0);
    const S lw = relu / (r2 + a.eta2);
    const S nu2 = a.nu_field ? S(2) * a.nu_field[f] : a.nu2;
    const S cv = -(nu2 * a.friction) * (volw * lw);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const S pp = (mf * mw) * (term * (df[d] * g));
      const S visc_f = cv * (S)(a.vel[f * D + d] - vw[d]);
      r[d].add(pp - mf * visc_f);
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) reaction[w * D + d] = r[d].get();
}

// ------------------------------------------------------------------ boundary-consistent density
// droplet_impact.py:1090-1112: the wall rows' mass scaled by rho_w / rho0
// (rho_w from the first wall pass) enters the fluid summation density; the
// fluid EOS follows, and the wall pass runs again on the corrected state.
template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_wall_mass_correction(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                                 const int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
  Sum<double> corr;
  for (int64_t e = off[i]; e < off[i + 1]; ++e) {
    const int64_t j = idx[e];
    if (j < a.n) continue;
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double df = sep(a, xi[d] - a.pos[j * D + d], d);
      r2 += df * df;
    }
    if (r2 > a.skip2d) continue;
    double w, g;
    cubic_wg<double>(sqrt(r2), a.kd, w, g);
    const double mw = (double)a.mass[j];
    corr.add((mw * ((double)a.rho[j] / a.rho0d) - mw) * w);
  }
  double rho_d = (double)a.rho[i] + hi(corr);
  rho_d = rho_d > (double)a.rho_floor ? rho_d : (double)a.rho_floor;
  const S rho = (S)rho_d;
  const S p = tait_hi<S>(rho_d, a.rho0d, a.bulk, a.gamma, a.ratio_floor);
  a.rho[i] = rho;
  a.p[i] = p;
  a.pr2[i] = p / (rho * rho);
  a.vol[i] = a.mass[i] / rho;
}

// ------------------------------------------------------------------ free-surface pressure closure
// droplet_impact.py:429-492, fluid-fluid pairs.  Rows below threshold * rho0
// receive p = max(0, sum_j w_j (p_j - g (y_i - y_j) rho_j) / sum_j w_j), with
// w_j = W_ij for every neighbour ("all", one pass) or only for neighbours
// whose pressure is known at the start of the pass ("interior", several
// passes).  Each pass reads (p_src, known_src) and writes (p_dst, known_dst);
// the last pass writes the state's p and p / rho^2.
template <typename S>
__global__ void k_closure_init(const S* __restrict__ rho, int64_t n, S thr, int* __restrict__ known) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) known[i] = rho[i] < thr ? 0 : 1;
}

template <int D, typename S>
__global__ void __launch_bounds__(kBlock) k_closure(WcsphArgs<S> a, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ idx,
                                                    const S* __restrict__ p_src, const int* __restrict__ k_src,
                                                    S* __restrict__ p_dst, int* __restrict__ k_dst,
                                                    int literal, int last) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  S pi = p_src[i];
  int ki = k_src[i];
  if (!ki) {
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = a.pos[i * D + d];
    Sum<S> num, den;
    for (int64_t e = off[i]; e < off[i + 1]; ++e) {
      const int64_t j = idx[e];
      if (j >= a.n) continue;
      S df[D], r2 = S(0);
#pragma unroll
      for (int d = 0; d < D; ++d) { df[d] = (S)sep(a, xi[d] - a.pos[j * D + d], d); r2 += df[d] * df[d]; }
      if (r2 > a.skip2) continue;
      if (!literal && !k_src[j]) continue;
      S w, g;
      cubic_wg<S>(sqrt_s<S>(r2), a.k, w, g);
      const S src = p_src[j] - (a.gravity * df[1]) * a.rho[j];
      num.add(w * src);
      den.add(w);
    }
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
    a.pr2[i] = pi / (r * r);
  }
}

// ------------------------------------------------------------------ host driver
template <int D, typename S>
int wcsph_rhs_impl(tsph_ctx* ctx, const tsph_wcsph_params* prm, const double* pos, const double* vel,
                   const void* mass, const double* wall_accel, const void* density_offset,
                   const void* nu_field, void* rho_out, void* p_out, void* acc_out,
                   void* wall_p, void* wall_rho, void* reaction, hipStream_t stream) {
  const int64_t n = prm->n_fluid, nw = prm->n_wall, nt = n + nw;
  TSPH_REQUIRE(ctx->nb_targets == n && ctx->wnb_targets == nw, TSPH_ERR_ARG,
               "neighbour lists were built for %lld fluid / %lld wall particles, rhs got %lld / %lld",
               (long long)ctx->nb_targets, (long long)ctx->wnb_targets, (long long)n, (long long)nw);
  TSPH_TRY(ctx->s_rho.ensure(sizeof(S) * (nt + 1)));
  TSPH_TRY(ctx->s_p.ensure(sizeof(S) * (nt + 1)));
  TSPH_TRY(ctx->s_aux.ensure(sizeof(S) * (nt + 1)));    // p / rho^2
  TSPH_TRY(ctx->s_aux2.ensure(sizeof(S) * (nt + 1)));   // m / rho
  TSPH_TRY(ctx->s_rho_raw.ensure(sizeof(S) * (nt + 1)));
  TSPH_TRY(ctx->s_gr.ensure(sizeof(S) * (n * D + 1)));
  WcsphArgs<S> a = make_wcsph_args<S>(*prm, D);
  a.pos = pos; a.vel = vel; a.mass = (const S*)mass; a.wall_accel = prm->has_wall_accel ? wall_accel : nullptr;
  a.rho = ctx->s_rho.as<S>(); a.p = ctx->s_p.as<S>(); a.pr2 = ctx->s_aux.as<S>(); a.vol = ctx->s_aux2.as<S>();
  a.dens_off = (const S*)density_offset;
  a.nu_field = (const S*)nu_field;
  const int64_t* off = ctx->nb_off.as<int64_t>();
  const int32_t* idx = ctx->nb_idx.as<int32_t>();
  if (n > 0) {
    const bool diff = prm->diffusion_scale > 0.0;
    if (!diff) {
      k_density<D, S><<<grid_xcd(n), kBlock, 0, stream>>>(a, off, idx, 1);
    } else {
      if (sizeof(S) == 4) {
        // low part of the raw FP32 density, so diffusion + EOS see it in FP64
        TSPH_TRY(ctx->s_aux3.ensure(sizeof(S) * (nt + 1)));
        a.rho_lo = ctx->s_aux3.as<S>();
      }
      WcsphArgs<S> a0 = a;
      a0.rho = ctx->s_rho_raw.as<S>();
      k_density<D, S><<<grid_xcd(n), kBlock, 0, stream>>>(a0, off, idx, 0);
      if (prm->diffusion_form == 1) {
        k_renorm_gradient<D, S><<<grid_for(n), kBlock, 0, stream>>>(a0, off, idx, ctx->s_gr.as<S>());
        k_diffusion<D, S, 1><<<grid_for(n), kBlock, 0, stream>>>(a, off, idx, a0.rho, ctx->s_gr.as<S>());
      } else {
        k_diffusion<D, S, 0><<<grid_for(n), kBlock, 0, stream>>>(a, off, idx, a0.rho, nullptr);
      }
    }
    TSPH_CHECK_LAUNCH();
    if (prm->surface_pressure_closure) {
      const int layers = prm->surface_pressure_donors_all ? 1
                         : (prm->surface_pressure_layers > 1 ? prm->surface_pressure_layers : 1);
      const size_t sb = (sizeof(S) * (size_t)(n + 1) + 255) / 256 * 256;
      const size_t kb = (sizeof(int) * (size_t)(n + 1) + 255) / 256 * 256;
      TSPH_TRY(ctx->s_cl.ensure(2 * sb + 2 * kb));
      char* base = ctx->s_cl.as<char>();
      S* pbuf[2] = {(S*)base, (S*)(base + sb)};
      int* kbuf[2] = {(int*)(base + 2 * sb), (int*)(base + 2 * sb + kb)};
      TSPH_CHECK_HIP(hipMemcpyAsync(pbuf[0], a.p, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
      k_closure_init<S><<<grid_for(n), kBlock, 0, stream>>>(
          a.rho, n, S(prm->surface_pressure_threshold * prm->rho0), kbuf[0]);
      for (int l = 0; l < layers; ++l) {
        const int src = l & 1, last = l == layers - 1;
        k_closure<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, off, idx, pbuf[src], kbuf[src],
                                                            last ? a.p : pbuf[src ^ 1], kbuf[src ^ 1],
                                                            prm->surface_pressure_donors_all, last);
      }
      TSPH_CHECK_LAUNCH();
    }
  }
  if (nw > 0) {
    k_wall_pressure<D, S><<<grid_for(nw), kBlock, 0, stream>>>(a, ctx->wnb_off.as<int64_t>(),
                                                               ctx->wnb_idx.as<int32_t>(),
                                                               (S*)wall_p, (S*)wall_rho);
    TSPH_CHECK_LAUNCH();
    if (prm->boundary_consistent_density && n > 0) {
      k_wall_mass_correction<D, S><<<grid_for(n), kBlock, 0, stream>>>(a, off, idx);
      k_wall_pressure<D, S><<<grid_for(nw), kBlock, 0, stream>>>(a, ctx->wnb_off.as<int64_t>(),
                                                                 ctx->wnb_idx.as<int32_t>(),
                                                                 (S*)wall_p, (S*)wall_rho);
      TSPH_CHECK_LAUNCH();
    }
  }
  if (n > 0) {
    k_force<D, S><<<grid_xcd(n), kBlock, 0, stream>>>(a, off, idx, (S*)acc_out);
    TSPH_CHECK_LAUNCH();
    TSPH_CHECK_HIP(hipMemcpyAsync(rho_out, a.rho, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
    TSPH_CHECK_HIP(hipMemcpyAsync(p_out, a.p, sizeof(S) * n, hipMemcpyDeviceToDevice, stream));
  }
  if (reaction && nw > 0 && prm->reaction_style != 0) {
    k_wall_loads<D, S><<<grid_for(nw), kBlock, 0, stream>>>(a, ctx->wnb_off.as<int64_t>(),
                                                            ctx->wnb_idx.as<int32_t>(), (S*)reaction);
    TSPH_CHECK_LAUNCH();
  }
  return TSPH_OK;
}


}  // namespace tsph

using namespace tsph;

extern "C" int tsph_wcsph_rhs(tsph_ctx* ctx, const tsph_wcsph_params* prm, const double* pos,
                              const double* vel, const void* mass, const double* wall_accel, void* rho,
                              void* p, void* acc, void* wall_p, void* wall_rho, void* reaction,
                              hipStream_t stream) {
  return tsph_wcsph_rhs_ex(ctx, prm, pos, vel, mass, wall_accel, nullptr, nullptr, rho, p, acc, wall_p,
                           wall_rho, reaction, stream);
}

extern "C" int tsph_wcsph_rhs_ex(tsph_ctx* ctx, const tsph_wcsph_params* prm, const double* pos,
                                 const double* vel, const void* mass, const double* wall_accel,
                                 const void* density_offset, const void* nu_field, void* rho, void* p,
                                 void* acc, void* wall_p, void* wall_rho, void* reaction, hipStream_t stream) {
  TSPH_REQUIRE(ctx && prm, TSPH_ERR_ARG, "ctx and params must not be NULL");
  TSPH_REQUIRE(prm->dim == 2 || prm->dim == 3, TSPH_ERR_ARG, "dim must be 2 or 3, got %d", prm->dim);
  // rhs_3d reads the surface-pressure closure (dam_break_3d.py:331-335) but none of the other
  // opt-in branches of the 2-D rhs
  const bool opts = prm->wrap_width != 0.0 || prm->density_renormalization || prm->boundary_consistent_density ||
                    density_offset || nu_field;
  TSPH_REQUIRE(!opts || prm->dim == 2, TSPH_ERR_ARG, "the opt-in rhs branches other than the closure are 2-D only");
  TSPH_REQUIRE(prm->wrap_width >= 0.0 && prm->wrap_width < INFINITY, TSPH_ERR_ARG,
               "wrap_width must be a positive finite length (0 disables)");
  TSPH_REQUIRE(!prm->density_renormalization || prm->lattice_kernel_sum > 0.0, TSPH_ERR_ARG,
               "density_renormalization needs a positive lattice_kernel_sum");
  TSPH_REQUIRE(!(density_offset && prm->density_renormalization), TSPH_ERR_ARG,
               "density_offset requires unnormalised summation density");
  TSPH_REQUIRE(prm->precision == TSPH_F64 || prm->precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(prm->h > 0.0, TSPH_ERR_ARG, "h must be positive");
  TSPH_REQUIRE(prm->n_fluid >= 0 && prm->n_wall >= 0, TSPH_ERR_ARG, "bad particle counts");
  TSPH_REQUIRE(prm->n_fluid == 0 || (pos && vel && mass && rho && p && acc), TSPH_ERR_ARG,
               "pos, vel, mass, rho, p and acc must not be NULL");
  if (prm->dim == 2) {
    if (prm->precision == TSPH_F64)
      return wcsph_rhs_impl<2, double>(ctx, prm, pos, vel, mass, wall_accel, density_offset, nu_field, rho, p, acc,
                                       wall_p, wall_rho, reaction, stream);
    return wcsph_rhs_impl<2, float>(ctx, prm, pos, vel, mass, wall_accel, density_offset, nu_field, rho, p, acc,
                                    wall_p, wall_rho, reaction, stream);
  }
  if (prm->precision == TSPH_F64)
    return wcsph_rhs_impl<3, double>(ctx, prm, pos, vel, mass, wall_accel, nullptr, nullptr, rho, p, acc, wall_p,
                                     wall_rho, reaction, stream);
  return wcsph_rhs_impl<3, float>(ctx, prm, pos, vel, mass, wall_accel, nullptr, nullptr, rho, p, acc, wall_p,
                                  wall_rho, reaction, stream);
}

// ------------------------------------------------------------------ elementwise helpers
namespace tsph {

template <typename S>
__global__ void k_tait(const S* __restrict__ rho, int64_t n, S rho0, S bulk, S gamma, S ratio_floor,
                       S* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = tait<S>(rho[i], rho0, bulk, gamma, ratio_floor);
}

template <typename S>
__global__ void k_kernel_eval(int kind, const S* __restrict__ r, int64_t n, KernelConst<S> k,
                              S* __restrict__ w, S* __restrict__ g, S* __restrict__ lap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  S ww, gg, ll = S(0);
  if (kind == 2) wendland_wg<S>(r[i], k, ww, gg);
  else if (kind == 1) cubic_wgl<S>(r[i], k, ww, gg, ll);
  else cubic_wg<S>(r[i], k, ww, gg);
  w[i] = ww;
  if (g) g[i] = gg;
  if (lap) lap[i] = ll;
}

}  // namespace tsph

extern "C" int tsph_eos_tait(const void* rho, int64_t n, int precision, double rho0, double sound_speed,
                             double gamma, double ratio_floor, void* p, hipStream_t stream) {
  TSPH_REQUIRE(n >= 0 && rho0 > 0.0 && gamma > 0.0, TSPH_ERR_ARG, "bad Tait parameters");
  if (n == 0) return TSPH_OK;
  const double bulk = rho0 * (sound_speed * sound_speed) / gamma;
  if (precision == TSPH_F64)
    k_tait<double><<<grid_for(n), kBlock, 0, stream>>>((const double*)rho, n, rho0, bulk, gamma,
                                                       ratio_floor, (double*)p);
  else
    k_tait<float><<<grid_for(n), kBlock, 0, stream>>>((const float*)rho, n, (float)rho0, (float)bulk,
                                                      (float)gamma, (float)ratio_floor, (float*)p);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

extern "C" int tsph_kernel_eval(int kind, const void* r, int64_t n, int precision, double h, void* w,
                                void* g, void* lap, hipStream_t stream) {
  TSPH_REQUIRE(h > 0.0 && n >= 0 && kind >= 0 && kind <= 3, TSPH_ERR_ARG, "bad kernel arguments");
  if (n == 0) return TSPH_OK;
  // kind: 0 cubic 2-D, 1 cubic 3-D (+ Laplacian), 2 Wendland C2 2-D
  if (precision == TSPH_F64) {
    KernelConst<double> k = kind == 2 ? make_wendland2d<double>(h) : make_cubic<double>(h, kind == 1 ? 3 : 2);
    k_kernel_eval<double><<<grid_for(n), kBlock, 0, stream>>>(kind, (const double*)r, n, k, (double*)w,
                                                              (double*)g, (double*)lap);
  } else {
    KernelConst<float> k = kind == 2 ? make_wendland2d<float>(h) : make_cubic<float>(h, kind == 1 ? 3 : 2);
    k_kernel_eval<float><<<grid_for(n), kBlock, 0, stream>>>(kind, (const float*)r, n, k, (float*)w,
                                                             (float*)g, (float*)lap);
  }
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

namespace tsph {
template <int D>
__global__ void k_apply_boundary(double* __restrict__ pos, double* __restrict__ vel, int64_t n, double lo0,
                                 double lo1, double lo2, double hi0, double hi1, double hi2, double tol,
                                 unsigned long long* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int hit_any = 0;
  if (i < n) {
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double x = pos[i * D + d];
      const int hit = (x < lo[d] - tol) || (x > hi[d] + tol);
      hit_any |= hit;
      pos[i * D + d] = x < lo[d] ? lo[d] : (x > hi[d] ? hi[d] : x);
      if (hit) vel[i * D + d] = 0.0;
    }
  }
  unsigned long long h = hit_any;
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
  if ((threadIdx.x & 63) == 0 && h) atomicAdd(count, h);
}
}  // namespace tsph

extern "C" int tsph_apply_boundary(double* pos, double* vel, int64_t n, int dim, const double* lo,
                                   const double* hi, double tol, unsigned long long* count,
                                   hipStream_t stream) {
  TSPH_REQUIRE(dim >= 1 && dim <= 3 && lo && hi && count, TSPH_ERR_ARG, "bad boundary arguments");
  TSPH_CHECK_HIP(hipMemsetAsync(count, 0, sizeof(unsigned long long), stream));
  if (n == 0) return TSPH_OK;
  const double l[3] = {lo[0], dim > 1 ? lo[1] : 0.0, dim > 2 ? lo[2] : 0.0};
  const double h[3] = {hi[0], dim > 1 ? hi[1] : 0.0, dim > 2 ? hi[2] : 0.0};
  const unsigned g = grid_for(n);
  if (dim == 1) k_apply_boundary<1><<<g, kBlock, 0, stream>>>(pos, vel, n, l[0], l[1], l[2], h[0], h[1], h[2], tol, count);
  else if (dim == 2) k_apply_boundary<2><<<g, kBlock, 0, stream>>>(pos, vel, n, l[0], l[1], l[2], h[0], h[1], h[2], tol, count);
  else k_apply_boundary<3><<<g, kBlock, 0, stream>>>(pos, vel, n, l[0], l[1], l[2], h[0], h[1], h[2], tol, count);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}

namespace tsph {
template <typename S>
__global__ void k_obstacle_clamp(double* __restrict__ pos, double* __restrict__ vel, const S* __restrict__ mass,
                                 int64_t n, double lo0, double lo1, double lo2, double hi0, double hi1,
                                 double hi2, double eps, unsigned long long* __restrict__ count,
                                 double* __restrict__ impulse) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int moved = 0;
  if (i < n) {
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    double x[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) { x[d] = pos[i * 3 + d]; v[d] = vel[i * 3 + d]; }
    const double m = mass ? (double)mass[i] : 0.0;
    moved = box_clamp(x, v, m, lo, hi, eps, mass ? impulse : nullptr);
    if (moved) {
#pragma unroll
      for (int d = 0; d < 3; ++d) { pos[i * 3 + d] = x[d]; vel[i * 3 + d] = v[d]; }
    }
  }
  unsigned long long c = moved;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}
}  // namespace tsph

extern "C" int tsph_obstacle_clamp(double* pos, double* vel, const void* mass, int precision, int64_t n,
                                   const double* lo, const double* hi, double eps, unsigned long long* count,
                                   double* impulse, hipStream_t stream) {
  TSPH_REQUIRE(lo && hi && count && n >= 0, TSPH_ERR_ARG, "bad obstacle clamp arguments");
  TSPH_REQUIRE(!mass || impulse, TSPH_ERR_ARG, "mass needs an impulse accumulator");
  TSPH_REQUIRE(precision == TSPH_F64 || precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_CHECK_HIP(hipMemsetAsync(count, 0, sizeof(unsigned long long), stream));
  if (n == 0) return TSPH_OK;
  TSPH_REQUIRE(pos && vel, TSPH_ERR_ARG, "pos and vel must not be NULL");
  const unsigned g = grid_for(n);
  if (precision == TSPH_F64)
    k_obstacle_clamp<double><<<g, kBlock, 0, stream>>>(pos, vel, (const double*)mass, n, lo[0], lo[1], lo[2],
                                                       hi[0], hi[1], hi[2], eps, count, impulse);
  else
    k_obstacle_clamp<float><<<g, kBlock, 0, stream>>>(pos, vel, (const float*)mass, n, lo[0], lo[1], lo[2],
                                                      hi[0], hi[1], hi[2], eps, count, impulse);
  TSPH_CHECK_LAUNCH();
  return TSPH_OK;
}
