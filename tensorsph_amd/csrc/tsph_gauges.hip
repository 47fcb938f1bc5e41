// Water-height gauges and obstacle pressure taps of the 3-D dam break against
// a box: kleefsman_obstacle_3d.py:273-318 (slab_height, peak_height,
// bypass_count) and marin_probes.py:25-79 (centreline_height,
// tap_node_pressures).  One pass over the fluid rows gives the row gauges in
// per-block partials (summed in a fixed order) and flags the rows within 2h
// of the instrumented faces; those rows are compacted in row order and one
// workgroup per tap quadrature node sums its one-sided SPH reconstruction
// over them.  Every sum has a fixed order, so the results do not depend on
// scheduling.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "tsph_common.hpp"
#include "tsph_engine.hpp"
#include "tsph_internal.hpp"

namespace tsph {

constexpr int kGaugeNodes = 8 * 7;
constexpr int kGaugeBlocks = 1024;

struct GaugeSpec {
  double sx0, sx1;            // slab: sx0 < x < sx1 and oz0 < z < oz1
  double oz0, oz1, ox1;       // obstacle span and downstream face
  double cx, cz;              // centreline station
  double inv_h, h2norm;       // 1 / h, 10 / (7 pi h^2)
  double near_lo[3], near_hi[3];
  double g;
  int per_tap, nodes;
  double node[kGaugeNodes][3];
};

// the 2-D cubic spline shape of marin_probes._cubic
TSPH_D double probe_cubic(double q) {
  if (q < 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
  const double t = 2.0 - q > 0.0 ? 2.0 - q : 0.0;
  return 0.25 * t * t * t;
}

enum { GR_SLAB = 0, GR_PEAK, GR_FRONT, GR_BYPASS, GR_H2, GR_N };

template <typename S>
__global__ void __launch_bounds__(256) k_gauge_rows(const double* __restrict__ pos, const S* __restrict__ vol,
                                                    int64_t n, GaugeSpec g, double* __restrict__ part,
                                                    unsigned char* __restrict__ flag) {
  double cnt = 0.0, byp = 0.0, h2 = 0.0, peak = -1.0e300, front = -1.0e300;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    if (x > g.sx0 && x < g.sx1 && z > g.oz0 && z < g.oz1) {
      cnt += 1.0;
      peak = y > peak ? y : peak;
    }
    front = x > front ? x : front;
    if (x > g.ox1 && (z < g.oz0 || z > g.oz1)) byp += 1.0;
    const double r = hypot(x - g.cx, z - g.cz);
    h2 += (double)vol[i] * probe_cubic(r * g.inv_h);
    flag[i] = (x >= g.near_lo[0] && x <= g.near_hi[0] && y <= g.near_hi[1] && z >= g.near_lo[2] &&
               z <= g.near_hi[2]) ? 1 : 0;
  }
  __shared__ double red[GR_N][256];
  red[GR_SLAB][threadIdx.x] = cnt;
  red[GR_PEAK][threadIdx.x] = peak;
  red[GR_FRONT][threadIdx.x] = front;
  red[GR_BYPASS][threadIdx.x] = byp;
  red[GR_H2][threadIdx.x] = h2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int o = threadIdx.x + s;
      red[GR_SLAB][threadIdx.x] += red[GR_SLAB][o];
      red[GR_BYPASS][threadIdx.x] += red[GR_BYPASS][o];
      red[GR_H2][threadIdx.x] += red[GR_H2][o];
      red[GR_PEAK][threadIdx.x] = fmax(red[GR_PEAK][threadIdx.x], red[GR_PEAK][o]);
      red[GR_FRONT][threadIdx.x] = fmax(red[GR_FRONT][threadIdx.x], red[GR_FRONT][o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < GR_N) part[(int64_t)blockIdx.x * GR_N + threadIdx.x] = red[threadIdx.x][0];
}

// one block of 256: partials of nblocks blocks, thread t takes t, t + 256, ...
__global__ void __launch_bounds__(256) k_gauge_reduce(const double* __restrict__ part, int nblocks,
                                                      double* __restrict__ out) {
  __shared__ double red[GR_N][256];
  double v[GR_N] = {0.0, -1.0e300, -1.0e300, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += 256) {
    const double* q = part + (int64_t)b * GR_N;
    v[GR_SLAB] += q[GR_SLAB];
    v[GR_BYPASS] += q[GR_BYPASS];
    v[GR_H2] += q[GR_H2];
    v[GR_PEAK] = fmax(v[GR_PEAK], q[GR_PEAK]);
    v[GR_FRONT] = fmax(v[GR_FRONT], q[GR_FRONT]);
  }
#pragma unroll
  for (int k = 0; k < GR_N; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int o = threadIdx.x + s;
      red[GR_SLAB][threadIdx.x] += red[GR_SLAB][o];
      red[GR_BYPASS][threadIdx.x] += red[GR_BYPASS][o];
      red[GR_H2][threadIdx.x] += red[GR_H2][o];
      red[GR_PEAK][threadIdx.x] = fmax(red[GR_PEAK][threadIdx.x], red[GR_PEAK][o]);
      red[GR_FRONT][threadIdx.x] = fmax(red[GR_FRONT][threadIdx.x], red[GR_FRONT][o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < GR_N) out[threadIdx.x] = red[threadIdx.x][0];
}

// One workgroup per quadrature node: sum over the rows near the faces of
// w = cubic(|x - node| / h) V (one-sided: upstream of the front face for taps
// 0-3, above the top face for taps 4-7) of w (p + rho g (y - y_node)), w and
// [w > 0].  out[3 node + 0..2].
template <typename S>
__global__ void __launch_bounds__(256) k_gauge_taps(const double* __restrict__ pos, const S* __restrict__ vol,
                                                    const S* __restrict__ rho, const S* __restrict__ p,
                                                    const int* __restrict__ idx, const int* __restrict__ count,
                                                    GaugeSpec g, double* __restrict__ out) {
  const int node = blockIdx.x;
  const int tap = node / g.per_tap;
  const double nx = g.node[node][0], ny = g.node[node][1], nz = g.node[node][2];
  const int m = *count;
  double num = 0.0, den = 0.0, sup = 0.0;
  for (int e = threadIdx.x; e < m; e += 256) {
    const int64_t i = idx[e];
    const double dx = pos[3 * i] - nx, dy = pos[3 * i + 1] - ny, dz = pos[3 * i + 2] - nz;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double w = probe_cubic(r * g.inv_h) * (double)vol[i];
    const bool side = tap < 4 ? dx <= 0.0 : dy >= 0.0;
    if (!side) w = 0.0;
    num += w * ((double)p[i] + (double)rho[i] * g.g * dy);
    den += w;
    sup += w > 0.0 ? 1.0 : 0.0;
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = num;
  red[1][threadIdx.x] = den;
  red[2][threadIdx.x] = sup;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * node + threadIdx.x] = red[threadIdx.x][0];
}

static int gauge_spec(const tsph_gauge_spec* s, GaugeSpec& g) {
  TSPH_REQUIRE(s->per_tap == 1 || s->per_tap == 7, TSPH_ERR_ARG, "per_tap must be 1 or 7, got %d", s->per_tap);
  TSPH_REQUIRE(s->h > 0.0 && s->aperture > 0.0, TSPH_ERR_ARG, "h and the gauge aperture must be positive");
  g = GaugeSpec{};
  g.sx0 = s->station_x - 0.5 * s->aperture;
  g.sx1 = s->station_x + 0.5 * s->aperture;
  g.oz0 = s->obstacle_lo[2];
  g.oz1 = s->obstacle_hi[2];
  g.ox1 = s->obstacle_hi[0];
  g.cx = s->station_x;
  g.cz = 0.5 * s->tank_depth;
  g.inv_h = 1.0 / s->h;
  g.h2norm = 10.0 / (7.0 * M_PI * s->h * s->h);
  g.g = s->gravity;
  g.per_tap = s->per_tap;
  g.nodes = 8 * s->per_tap;
  double zmin = 1e300, zmax = -1e300;
  for (int k = 0; k < g.nodes; ++k) {
    for (int d = 0; d < 3; ++d) g.node[k][d] = s->nodes[3 * k + d];
    zmin = g.node[k][2] < zmin ? g.node[k][2] : zmin;
    zmax = g.node[k][2] > zmax ? g.node[k][2] : zmax;
  }
  // rows beyond 2h of every node carry no weight, so the prefilter changes no sum
  const double r = 2.0 * s->h;
  g.near_lo[0] = s->obstacle_lo[0] - r;
  g.near_hi[0] = s->obstacle_hi[0] + r;
  g.near_lo[1] = -1e300;
  g.near_hi[1] = s->obstacle_hi[1] + r;
  g.near_lo[2] = zmin - r;
  g.near_hi[2] = zmax + r;
  return TSPH_OK;
}

template <typename S>
int gauges_impl(tsph_ctx* ctx, const GaugeSpec& g, const tsph_gauge_spec* spec, int64_t n, const double* pos,
                const S* vol, const S* rho, const S* p, double* out, hipStream_t stream) {
  TSPH_REQUIRE(n < (1ll << 31) - 1, TSPH_ERR_ARG, "too many rows for the gauge compaction");
  const int G = n > 0 ? (int)std::min<int64_t>(grid_for(n), kGaugeBlocks) : 1;
  const int nn = n > 0 ? (int)n : 1;
  TSPH_TRY(ctx->g_part.ensure(sizeof(double) * GR_N * G));
  TSPH_TRY(ctx->g_flag.ensure((size_t)nn));
  TSPH_TRY(ctx->g_idx.ensure(sizeof(int) * (size_t)nn + 256));
  TSPH_TRY(ctx->g_out.ensure(sizeof(double) * (GR_N + 3 * kGaugeNodes)));
  double* dout = ctx->g_out.as<double>();
  int* cnt = (int*)(ctx->g_idx.as<char>() + sizeof(int) * (size_t)nn);
  TSPH_CHECK_HIP(hipMemsetAsync(dout, 0, sizeof(double) * (GR_N + 3 * kGaugeNodes), stream));
  if (n > 0) {
    k_gauge_rows<S><<<G, 256, 0, stream>>>(pos, vol, n, g, ctx->g_part.as<double>(), ctx->g_flag.as<unsigned char>());
    k_gauge_reduce<<<1, 256, 0, stream>>>(ctx->g_part.as<double>(), G, dout);
    TSPH_CHECK_LAUNCH();
    size_t bytes = 0;
    hipcub::CountingInputIterator<int> rows(0);
    TSPH_CHECK_HIP(hipcub::DeviceSelect::Flagged(nullptr, bytes, rows, ctx->g_flag.as<unsigned char>(),
                                                 ctx->g_idx.as<int>(), cnt, (int)n, stream));
    TSPH_TRY(ctx->g_tmp.ensure(bytes > 0 ? bytes : 1));
    TSPH_CHECK_HIP(hipcub::DeviceSelect::Flagged(ctx->g_tmp.ptr, bytes, rows, ctx->g_flag.as<unsigned char>(),
                                                 ctx->g_idx.as<int>(), cnt, (int)n, stream));
    k_gauge_taps<S><<<g.nodes, 256, 0, stream>>>(pos, vol, rho, p, ctx->g_idx.as<int>(), cnt, g, dout + GR_N);
    TSPH_CHECK_LAUNCH();
  }
  double h[GR_N + 3 * kGaugeNodes];
  TSPH_CHECK_HIP(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, stream));
  TSPH_CHECK_HIP(hipStreamSynchronize(stream));
  const double count = h[GR_SLAB];
  out[0] = count * spec->dx * spec->dx * spec->dx / (spec->aperture * (g.oz1 - g.oz0));
  out[1] = count > 0.0 ? h[GR_PEAK] : 0.0;
  out[2] = n > 0 ? h[GR_FRONT] : 0.0;
  out[3] = h[GR_BYPASS];
  out[4] = h[GR_H2] * g.h2norm;
  const double tiny = sizeof(S) == 8 ? 2.2250738585072014e-308 : 1.1754943508222875e-38;
  for (int k = 0; k < kGaugeNodes; ++k) {
    double pk = 0.0, sk = 0.0;
    if (k < g.nodes) {
      const double num = h[GR_N + 3 * k], den = h[GR_N + 3 * k + 1];
      pk = num / (den > tiny ? den : tiny);
      pk = pk > 0.0 ? pk : 0.0;
      sk = h[GR_N + 3 * k + 2];
    }
    out[5 + 2 * k] = pk;
    out[6 + 2 * k] = sk;
  }
  return TSPH_OK;
}

}  // namespace tsph

using namespace tsph;

extern "C" int tsph_gauges(tsph_ctx* ctx, const tsph_gauge_spec* spec, int precision, int64_t n,
                           const double* pos, const void* vol, const void* rho, const void* p, double* out,
                           hipStream_t stream) {
  TSPH_REQUIRE(ctx && spec && out && n >= 0, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(precision == TSPH_F64 || precision == TSPH_F32, TSPH_ERR_ARG, "bad precision");
  TSPH_REQUIRE(n == 0 || (pos && vol && rho && p), TSPH_ERR_ARG, "NULL array");
  GaugeSpec g;
  TSPH_TRY(gauge_spec(spec, g));
  if (precision == TSPH_F64)
    return gauges_impl<double>(ctx, g, spec, n, pos, (const double*)vol, (const double*)rho, (const double*)p, out,
                               stream);
  return gauges_impl<float>(ctx, g, spec, n, pos, (const float*)vol, (const float*)rho, (const float*)p, out,
                            stream);
}

extern "C" int tsph_sim_gauges(tsph_sim* sim, const tsph_gauge_spec* spec, double* out, hipStream_t stream) {
  TSPH_REQUIRE(sim && spec && out, TSPH_ERR_ARG, "bad arguments");
  TSPH_REQUIRE(sim->dim == 3, TSPH_ERR_ARG, "the gauges are defined for the 3-D stepper");
  TSPH_REQUIRE(sim->steps_done > 0, TSPH_ERR_ARG, "no step has been evaluated yet");
  GaugeSpec g;
  TSPH_TRY(gauge_spec(spec, g));
  // the input side of the last step: its positions and the density / pressure
  // / volume its evaluation produced, all in the stepper's row order
  const double* pos = sim->pos[sim->cur ^ 1].as<double>();
  if (sim->prec == TSPH_F64)
    return gauges_impl<double>(sim->ctx, g, spec, sim->n, pos, sim->vol.as<double>(), sim->rho.as<double>(),
                               sim->p.as<double>(), out, stream);
  return gauges_impl<float>(sim->ctx, g, spec, sim->n, pos, sim->vol.as<float>(), sim->rho.as<float>(),
                            sim->p.as<float>(), out, stream);
}
