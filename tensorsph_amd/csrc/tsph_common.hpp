// Shared device helpers for the tensorsph_amd particle pipeline (gfx950).
//
// Conventions used by every kernel:
//   * particle coordinates and velocities are FP64, array-of-structs (N, D),
//     the layout of the (N, D) tensors the Python API hands over;
//   * per-particle scalars (mass, density, pressure) and pair arithmetic use
//     the precision type S: double ("plain FP64") or float ("FP32
//     compensated": pair deltas formed in FP64 then rounded, kernel math in
//     FP32, sums Kahan-compensated);
//   * every per-particle sum is a gather over that particle's own neighbour
//     list in a fixed order, so results never depend on thread scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TSPH_HD __host__ __device__ __forceinline__
#define TSPH_D __device__ __forceinline__

namespace tsph {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- accumulators
template <typename S> struct Sum;
// add_if(keep, x) is add(x) when keep and leaves the sum untouched otherwise,
// so a sweep can evaluate every candidate without branching and still sum
// exactly the terms the branching form sums.
template <> struct Sum<double> {
  double s = 0.0;
  TSPH_D void add(double x) { s += x; }
  TSPH_D void add_if(bool keep, double x) { s = keep ? s + x : s; }
  // four terms in order, one after the other (the order of add)
  TSPH_D void add4(double x0, double x1, double x2, double x3) { s = (((s + x0) + x1) + x2) + x3; }
  TSPH_D double get() const { return s; }
};
// Kahan summation: the running compensation c carries the low-order bits that
// s + x drops.  Built without fast-math, so the compiler keeps the order.
template <> struct Sum<float> {
  float s = 0.0f, c = 0.0f;
  TSPH_D void add(float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  TSPH_D void add_if(bool keep, float x) {
    const float y = x - c;
    const float t = s + y;
    const float cn = (t - s) - y;
    c = keep ? cn : c;
    s = keep ? t : s;
  }
  // the four terms of one neighbour chunk summed pairwise, then one
  // compensated addition: the rounding of a four-term partial is that of its
  // terms, and the compensation carries every partial the running sum drops
  TSPH_D void add4(float x0, float x1, float x2, float x3) { add((x0 + x1) + (x2 + x3)); }
  TSPH_D float get() const { return s; }
};

// Value of an accumulator to FP64 accuracy.  For the Kahan sum that is s - c:
// c holds the negative of the low-order bits the running sum s dropped.
TSPH_D double hi(const Sum<double>& a) { return a.s; }
TSPH_D double hi(const Sum<float>& a) { return (double)a.s - (double)a.c; }

// Four values of one particle that a pair sweep reads together, so that a
// neighbour costs one 16-byte (FP32) or two 16-byte (FP64) gathers.
template <typename S> struct alignas(4 * sizeof(S)) Vec4 {
  S x, y, z, w;
};

template <int D, typename S> struct VSum {
  Sum<S> v[D];
  TSPH_D void add(const S* x) {
#pragma unroll
    for (int d = 0; d < D; ++d) v[d].add(x[d]);
  }
};

// ---------------------------------------------------------------- kernels
// Cubic spline (Monaghan 1992), q = r/h, support 2h.
//   2-D sigma = 10/(7 pi h^2)   (droplet_impact.py:276-304)
//   3-D sigma = 1/(pi h^3)      (threed/kernel3.py:48-123)
// Returned: W, the gradient scale g with grad_i W = delta * g, and (3-D) the
// radial Laplacian f'' + 2 f'/q with its q -> 0 limit -6.
template <typename S> struct KernelConst {
  S h, inv_h, sigma, sigma_over_h, w0;
  double w0d;   // W(0) in FP64, for the self term of the density
};

template <typename S> TSPH_D void cubic_wg(S r, const KernelConst<S>& k, S& w, S& g) {
  const S q = r * k.inv_h;
  S shape, dshape;
  if (q < S(1)) {
    shape = S(1) - S(1.5) * (q * q) + S(0.75) * (q * q * q);
    dshape = S(-3) * q + S(2.25) * (q * q);
  } else if (q < S(2)) {
    const S t = S(2) - q;
    shape = S(0.25) * (t * t * t);
    dshape = S(-0.75) * (t * t);
  } else {
    shape = S(0);
    dshape = S(0);
  }
  w = k.sigma * (shape > S(0) ? shape : S(0));
  const S rc = r > S(1e-12) ? r : S(1e-12);
  g = k.sigma_over_h * dshape / rc;
}

// cubic_wg's W and g, the same expressions, with the pieces selected instead
// of branched on: a pair sweep's entries then form straight-line code whose
// loads can all be issued ahead of the arithmetic.
template <typename S> TSPH_D void cubic_wg_select(S r, const KernelConst<S>& k, S& w, S& g) {
  const S q = r * k.inv_h;
  const S shape1 = S(1) - S(1.5) * (q * q) + S(0.75) * (q * q * q);
  const S dshape1 = S(-3) * q + S(2.25) * (q * q);
  const S t = S(2) - q;
  const S shape2 = S(0.25) * (t * t * t);
  const S dshape2 = S(-0.75) * (t * t);
  const bool in1 = q < S(1), in2 = q < S(2);
  const S shape = in1 ? shape1 : (in2 ? shape2 : S(0));
  const S dshape = in1 ? dshape1 : (in2 ? dshape2 : S(0));
  w = k.sigma * (shape > S(0) ? shape : S(0));
  const S rc = r > S(1e-12) ? r : S(1e-12);
  g = k.sigma_over_h * dshape / rc;
}

// Gradient factor g of the cubic spline (grad_i W = g (x_i - x_j)) from the
// squared separation.  FP64 is cubic_wg's exact square root and quotient.
// FP32 takes r and 1/r from one hardware reciprocal square root (v_rsq_f32,
// 1 ulp), which removes the square root and the division of the piecewise
// form; r is clamped at 1e-12 as in cubic_wg.
template <typename S> TSPH_D S cubic_g_r2(S r2, const KernelConst<S>& k);
template <> TSPH_D double cubic_g_r2<double>(double r2, const KernelConst<double>& k) {
  double w, g;
  cubic_wg_select<double>(sqrt(r2), k, w, g);
  return g;
}
template <> TSPH_D float cubic_g_r2<float>(float r2, const KernelConst<float>& k) {
  const float rinv = r2 > 1e-24f ? __builtin_amdgcn_rsqf(r2) : 1e12f;
  const float q = (r2 * rinv) * k.inv_h;
  // both pieces with sigma / h folded in, then one select: 2 - q clamped at
  // 0 makes the outer piece vanish beyond the support (a nested select here
  // compiles to a branch)
  const float s = k.sigma_over_h;
  const float d1 = q * fmaf(2.25f * s, q, -3.0f * s);
  const float t = fmaxf(2.0f - q, 0.0f);
  const float d2 = (-0.75f * s) * (t * t);
  return (q < 1.0f ? d1 : d2) * rinv;
}

// sqrt(r2) for a squared separation, to about 5e-14 relative: the FP32
// hardware reciprocal square root (v_rsq_f32, relative error below 2e-7 after
// rounding r2 to FP32) refined by one Newton step in FP64, which squares the
// error.  Separations below 1e-15 return 0.  The FP32-compensated density
// sweeps use it in place of the correctly rounded FP64 square root.
TSPH_D double sqrt_r2_newton(double r2) {
  const double y0 = (double)__builtin_amdgcn_rsqf((float)r2);
  const double y1 = y0 * (1.5 - (0.5 * r2) * (y0 * y0));
  return r2 > 1.0e-30 ? r2 * y1 : 0.0;
}

// W of the cubic spline for r known to FP64 accuracy, branch-free: with
// a = (2 - q)+ and b = (1 - q)+ both pieces are (a^3 - 4 b^3) / 4 (expanding
// (2 - q)^3 - 4 (1 - q)^3 gives 4 - 6 q^2 + 3 q^3 on q < 1), so no select.
// Agrees with cubic_wg's W to FP64 rounding; the FP32-compensated wall sweep
// uses it.
TSPH_D double cubic_w_select(double r, const KernelConst<double>& k) {
  const double q = r * k.inv_h;
  const double a = fmax(2.0 - q, 0.0), b = fmax(1.0 - q, 0.0);
  return (0.25 * k.sigma) * fma(a * a, a, -4.0 * (b * b * b));
}

// W and dW/dr of the cubic spline for r known to FP64 accuracy, branch-free:
// cubic_w_select's W and the derivative cubic_wg_select forms before its
// division by r (the same values to FP64 rounding).
TSPH_D void cubic_w_dw_select(double r, const KernelConst<double>& k, double& w, double& dw) {
  // with a = (2 - q)+ and b = (1 - q)+ both pieces are one expression:
  // shape = (a^3 - 4 b^3) / 4 and dshape = 3 b^2 - 3 a^2 / 4 (expanding
  // (2 - q)^3 - 4 (1 - q)^3 gives 4 - 6 q^2 + 3 q^3 on q < 1), so neither
  // value needs a select
  const double q = r * k.inv_h;
  const double a = fmax(2.0 - q, 0.0), b = fmax(1.0 - q, 0.0);
  const double a2 = a * a, b2 = b * b;
  w = (0.25 * k.sigma) * fma(a2, a, -4.0 * (b2 * b));
  dw = k.sigma_over_h * fma(3.0, b2, -0.75 * a2);
}

// Byte offset into an array with a 32-bit index, so the address is the
// array's (uniform) base plus a zero-extended 32-bit offset.
template <typename T> TSPH_D const T* at_bytes(const T* base, uint32_t bytes) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + bytes);
}

// a / b; FP32 as a times the hardware reciprocal of b (v_rcp_f32, 1 ulp).
template <typename S> TSPH_D S quot(S a, S b);
template <> TSPH_D double quot<double>(double a, double b) { return a / b; }
template <> TSPH_D float quot<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename S> TSPH_D void cubic_wgl(S r, const KernelConst<S>& k, S& w, S& g, S& lap) {
  const S q = r * k.inv_h;
  S shape, fp, fpp;
  if (q < S(1)) {
    shape = S(1) - S(1.5) * (q * q) + S(0.75) * (q * q * q);
    fp = S(-3) * q + S(2.25) * (q * q);
    fpp = S(-3) + S(4.5) * q;
  } else if (q < S(2)) {
    const S t = S(2) - q;
    shape = S(0.25) * (t * t * t);
    fp = S(-0.75) * (t * t);
    fpp = S(1.5) * t;
  } else {
    shape = S(0); fp = S(0); fpp = S(0);
  }
  w = k.sigma * (shape > S(0) ? shape : S(0));
  const S rc = r > S(1e-12) ? r : S(1e-12);
  g = k.sigma_over_h * fp / rc;
  const S qc = q > S(1e-12) ? q : S(1e-12);
  const S twice = q > S(1e-10) ? S(2) * fp / qc : S(-6);
  lap = q < S(2) ? k.sigma * (fpp + twice) * (k.inv_h * k.inv_h) : S(0);
}

// Gradient factor g and Laplacian lap of the cubic spline from the squared
// separation (cubic_wgl's values).  FP64 is cubic_wgl itself; FP32 takes r
// and 1/r from one hardware reciprocal square root, so 2 f'(q) / q becomes
// 2 f'(q) h / r without a division.
template <typename S> TSPH_D void cubic_gl_r2(S r2, const KernelConst<S>& k, S& g, S& lap);
template <> TSPH_D void cubic_gl_r2<double>(double r2, const KernelConst<double>& k, double& g, double& lap) {
  double w;
  cubic_wgl<double>(sqrt(r2), k, w, g, lap);
}
template <> TSPH_D void cubic_gl_r2<float>(float r2, const KernelConst<float>& k, float& g, float& lap) {
  const float rinv = r2 > 1e-24f ? __builtin_amdgcn_rsqf(r2) : 1e12f;
  const float q = (r2 * rinv) * k.inv_h;
  float fp, fpp;
  if (q < 1.0f) {
    fp = -3.0f * q + 2.25f * (q * q);
    fpp = -3.0f + 4.5f * q;
  } else if (q < 2.0f) {
    const float t = 2.0f - q;
    fp = -0.75f * (t * t);
    fpp = 1.5f * t;
  } else {
    fp = 0.0f; fpp = 0.0f;
  }
  g = k.sigma_over_h * fp * rinv;
  const float twice = q > 1e-10f ? 2.0f * fp * (k.h * rinv) : -6.0f;
  lap = q < 2.0f ? k.sigma * (fpp + twice) * (k.inv_h * k.inv_h) : 0.0f;
}

// Wendland C2, 2-D sigma = 7/(4 pi h^2) (droplet_impact.py:320-346).
template <typename S> TSPH_D void wendland_wg(S r, const KernelConst<S>& k, S& w, S& g) {
  const S q = r * k.inv_h;
  S t = S(1) - S(0.5) * q;
  t = t > S(0) ? t : S(0);
  const S t2 = t * t;
  w = k.sigma * (t2 * t2) * (S(2) * q + S(1));
  const S rc = r > S(1e-12) ? r : S(1e-12);
  g = k.sigma_over_h * (S(-5) * q * (t2 * t)) / rc;
}

template <typename S>
inline KernelConst<S> make_cubic(double h, int dim) {
  KernelConst<S> k;
  double sigma = dim == 3 ? 1.0 / (kPi * h * h * h) : 10.0 / (7.0 * kPi * h * h);
  k.h = S(h); k.inv_h = S(1.0 / h); k.sigma = S(sigma); k.sigma_over_h = S(sigma / h);
  k.w0 = S(sigma);
  k.w0d = sigma;
  return k;
}

template <typename S>
inline KernelConst<S> make_wendland2d(double h) {
  KernelConst<S> k;
  double sigma = 7.0 / (4.0 * kPi * h * h);
  k.h = S(h); k.inv_h = S(1.0 / h); k.sigma = S(sigma); k.sigma_over_h = S(sigma / h);
  k.w0 = S(sigma);
  k.w0d = sigma;
  return k;
}

// ---------------------------------------------------------------- small math
template <typename S> TSPH_D S sqrt_s(S x);
template <> TSPH_D double sqrt_s<double>(double x) { return sqrt(x); }
template <> TSPH_D float sqrt_s<float>(float x) { return sqrtf(x); }
template <typename S> TSPH_D S pow_s(S x, S y);
template <> TSPH_D double pow_s<double>(double x, double y) { return pow(x, y); }
template <> TSPH_D float pow_s<float>(float x, float y) { return powf(x, y); }

// Tait EOS of droplet_impact.tait_pressure (:415-427): ratio floored at 0.01,
// pressure clamped at zero.  ``ratio_floor`` < 0 gives the 3-D form
// (threed/operators3.py:82-86) without the floor.
template <typename S>
TSPH_D S tait(S rho, S rho0, S bulk, S gamma, S ratio_floor) {
  S ratio = rho / rho0;
  if (ratio_floor > S(0) && ratio < ratio_floor) ratio = ratio_floor;
  S p = bulk * (pow_s<S>(ratio, gamma) - S(1));
  return p > S(0) ? p : S(0);
}

// Tait EOS of a density known to FP64 accuracy (rho from hi() plus the self
// term).  FP64: the power form above.  FP32: eps = ratio - 1 is formed in FP64
// and p = B expm1(gamma log1p(eps)) is evaluated in FP32, which is accurate
// relative to p itself; forming ratio^gamma - 1 from an FP32-rounded density
// instead cancels to an error of about gamma B 2^-24, far above p near rho0.
template <typename S>
TSPH_D S tait_hi(double rho, double rho0, S bulk, S gamma, S ratio_floor);
template <>
TSPH_D double tait_hi<double>(double rho, double rho0, double bulk, double gamma, double ratio_floor) {
  return tait<double>(rho, rho0, bulk, gamma, ratio_floor);
}
template <>
TSPH_D float tait_hi<float>(double rho, double rho0, float bulk, float gamma, float ratio_floor) {
  double ratio = rho / rho0;
  if (ratio_floor > 0.0f && ratio < (double)ratio_floor) ratio = (double)ratio_floor;
  const float p = bulk * expm1f(gamma * log1pf((float)(ratio - 1.0)));
  return p > 0.0f ? p : 0.0f;
}

// Stores an FP64 density as S and, when lo != NULL, the part S dropped, so a
// later sweep can rebuild the FP64 value (lo[i] = 0 in FP64 mode).
template <typename S>
TSPH_D S store_rho(double rho_d, S* rho, S* lo, int64_t i) {
  const S r = (S)rho_d;
  rho[i] = r;
  if (lo) lo[i] = (S)(rho_d - (double)r);
  return r;
}

// ---------------------------------------------------------------- launch helpers
constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// Chunked block -> work mapping for the 8 XCDs of MI355X.  Consecutive
// workgroup ids are dispatched round-robin over the XCDs; remapping so XCD x
// gets the x-th contiguous chunk of blocks keeps spatially adjacent particles
// (and so their shared neighbours) in one XCD's L2.  The grid is padded to a
// multiple of 8 (grid_xcd), which makes the map a bijection; kernels bound-
// check the particle index as usual.
TSPH_D int64_t xcd_block(int64_t bid, int64_t nblocks_padded) {
  const int64_t per = nblocks_padded >> 3;
  return (bid & 7) * per + (bid >> 3);
}
inline unsigned grid_xcd(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  g = ((g + 7) / 8) * 8;
  return (unsigned)(g > 0 ? g : 8);
}
// Row block of the calling workgroup: xcd_block on a grid padded by grid_xcd
// (neighbouring row blocks share one XCD's L2), the plain block index on any
// other grid.
TSPH_D int64_t row_block() {
  const int64_t nb = gridDim.x;
  return (nb & 7) ? (int64_t)blockIdx.x : xcd_block(blockIdx.x, nb);
}

}  // namespace tsph
