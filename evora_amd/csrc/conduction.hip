// conduction.hip -- thermal conduction (conduction.f90:26-164) and its time
// step limit (timestep.f90:72-104) on the device.
//
// k_conduct is one HBM pass: a block owns a 32 x 8 column tile and marches
// a chunk of z planes.  Per plane it forms the temperature T = p / rho *
// temp0 (comp.f90:96-100, no chemistry) of the tile and of its one-cell
// ring in LDS, with the ghost cells of `bound` folded into the load index;
// z neighbours stay in registers.  Each interior cell then takes the x, y
// and z temperature-flux differences, in the reference's order, into q(5)
// and q(6) (thermalintegration), writes its new q to the other buffer, and
// feeds the CFL limit of `hydrostep` and the conduction limit of
// `thermalstep` of the new state into the status block, so the next step
// needs no pass of its own.  Only simplekappa: kappa_eff = 1/g1.
//
// The arithmetic is the reference's (this unit is compiled without FMA
// contraction); the kernel is HBM-bound, so IEEE division costs nothing
// measurable.  k_thermal_min gives the conduction limit of a state that did
// not come out of k_conduct (after an upload or a sweep).
#include "evh_internal.h"

namespace evh {
namespace {

constexpr int CX = 32, CY = 8, CT = CX * CY;  // tile and threads
constexpr int RX = CX + 2, RY = CY + 2;        // with the one-cell ring
constexpr int NRING = 2 * CX + 2 * CY;         // ring cells (no corners)

struct CGeo {
  int nx, ny, nz;
  long long n;
  int zchunk, tiles_x, tiles_y, nchunks, ntiles;
};

// temperature of a cell from its conserved q (pressure, comp.f90:45-85)
__device__ __forceinline__ double cell_temp(const KParams &P, double q[6], int &err) {
  double w[5];
  err = q2w(P, q, w);
  return w[4] / w[0] * P.temp0;
}

template <typename S>
__device__ __forceinline__ void load_q(const CGeo &G, const S *__restrict__ q, const KParams &P, int gx, int gy,
                                       int gz, double out[6]) {
  // ghost cells of `bound` (boundary.f90): mirror or periodic source cell;
  // the upper-x reflection of boundtype 3 negates velocities only, which T
  // does not see.  Indices beyond the ghost layer (partial tiles) are
  // clamped: those lanes update nothing.
  const bool per = P.boundtype == 2;
  gx = min(max(gx, -2), G.nx + 1);
  gy = min(max(gy, -2), G.ny + 1);
  gz = min(max(gz, -2), G.nz + 1);
  const int sx = ghost_src(gx, G.nx, per), sy = ghost_src(gy, G.ny, per), sz = ghost_src(gz, G.nz, per);
  const long long c = (long long)sx + (long long)G.nx * ((long long)sy + (long long)G.ny * sz);
#pragma unroll
  for (int m = 0; m < 6; m++) out[m] = double(q[m * G.n + c]);
}

// thermalflux with simplekappa (conduction.f90:126-144)
__device__ __forceinline__ double tflux(const KParams &P, double tl, double tr, double dinv) {
  const double kappa_eff = 1.0 / P.g1;
  const double dtempdx = (tr - tl) * dinv;
  return -kappa_eff * dtempdx;
}

// thermalintegration (conduction.f90:146-162); pg = q(1)**g1
__device__ __forceinline__ void tintegrate(const KParams &P, double jl, double jr, double q[6], double delta,
                                           double pg) {
  const double de = delta * (jl - jr);
  q[4] = q[4] + de;
  q[5] = q[5] + P.g1 * de / pg;
}

// min over cells of rho / kappa, kappa = kappa0 T**2.5 (timestep.f90:84-93)
__device__ __forceinline__ double tstep_cell(const KParams &P, double rho, double temp) {
  return rho / (P.kappa0 * pow(temp, 2.5));
}

constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int j = b % kXcds, i = b / kXcds;
  const int base = ntiles / kXcds, rem = ntiles % kXcds;
  const int count = base + (j < rem ? 1 : 0);
  if (i >= count) return -1;
  return j * base + (j < rem ? j : rem) + i;
}

template <typename S>
__global__ void __launch_bounds__(CT) k_conduct(KParams P, CGeo G, const S *__restrict__ qin, S *__restrict__ qout,
                                                double dt, Status *st) {
  __shared__ double T[RY * RX];
  const int lin = xcd_tile(blockIdx.x, G.ntiles);
  if (lin < 0) return;  // padding block (whole block exits before any barrier)
  const int tile = lin % (G.tiles_x * G.tiles_y), chunk = lin / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * CX, y0 = (tile / G.tiles_x) * CY;
  const int z0 = chunk * G.zchunk, z1 = min(z0 + G.zchunk, G.nz);
  const int tid = threadIdx.x, cx = tid % CX, cy = tid / CX;
  const int gx = x0 + cx, gy = y0 + cy;
  const bool inside = gx < G.nx && gy < G.ny;
  // ring job: lanes 0..NRING-1 each own one ring cell of the plane
  int rx = 0, ry = 0;
  const bool ring = tid < NRING;
  if (tid < CX) { rx = tid; ry = -1; }
  else if (tid < 2 * CX) { rx = tid - CX; ry = CY; }
  else if (tid < 2 * CX + CY) { rx = -1; ry = tid - 2 * CX; }
  else if (tid < NRING) { rx = CX; ry = tid - 2 * CX - CY; }

  const double dx_delta = dt * P.dxinv, dy_delta = dt * P.dyinv, dz_delta = dt * P.dzinv;
  double vmin = 1e6, tvmin = 1e99;

  // qc: q of plane k, qn of plane k+1; tm, tc, tn their temperatures.
  // (cell_temp rewrites q only with dual energy, which conduction refuses.)
  double qc[6], qn[6], tmp[6];
  int e;
  load_q(G, qin, P, gx, gy, z0 - 1, tmp);
  double tm = cell_temp(P, tmp, e);
  load_q(G, qin, P, gx, gy, z0, qc);
#pragma unroll
  for (int m = 0; m < 6; m++) tmp[m] = qc[m];
  double tc = cell_temp(P, tmp, e);
  for (int k = z0; k < z1; k++) {
    // the primitive of main.f90:171 before thermalconduction stops on a
    // non-positive pressure of this cell
    if (e && inside) report(st, e, (unsigned long long)(gx + (long long)G.nx * (gy + (long long)G.ny * k)));
    load_q(G, qin, P, gx, gy, k + 1, qn);
#pragma unroll
    for (int m = 0; m < 6; m++) tmp[m] = qn[m];
    int en;
    const double tn = cell_temp(P, tmp, en);
    T[(cx + 1) + RX * (cy + 1)] = tc;
    if (ring) {
      double qr[6];
      load_q(G, qin, P, x0 + rx, y0 + ry, k, qr);
      int er;
      T[(rx + 1) + RX * (ry + 1)] = cell_temp(P, qr, er);
    }
    __syncthreads();
    if (inside) {
      const int me = (cx + 1) + RX * (cy + 1);
      const double pg = pow(qc[0], P.g1);
      double jl = tflux(P, T[me - 1], tc, P.dxinv), jr = tflux(P, tc, T[me + 1], P.dxinv);
      if (G.nx != 1) tintegrate(P, jl, jr, qc, dx_delta, pg);
      jl = tflux(P, T[me - RX], tc, P.dyinv);
      jr = tflux(P, tc, T[me + RX], P.dyinv);
      if (G.ny != 1) tintegrate(P, jl, jr, qc, dy_delta, pg);
      jl = tflux(P, tm, tc, P.dzinv);
      jr = tflux(P, tc, tn, P.dzinv);
      if (G.nz != 1) tintegrate(P, jl, jr, qc, dz_delta, pg);
      const long long c = (long long)gx + (long long)G.nx * ((long long)gy + (long long)G.ny * k);
#pragma unroll
      for (int m = 0; m < 6; m++) qout[m * G.n + c] = S(qc[m]);
      // the next step's limits from the new state (as stored)
      double qs[6], w[5];
#pragma unroll
      for (int m = 0; m < 6; m++) qs[m] = double(S(qc[m]));
      const int ep = q2w(P, qs, w);
      if (ep) report_next(st, ep, (unsigned long long)c);
      vmin = dmin(vmin, cfl_cell(P, w));
      tvmin = dmin(tvmin, tstep_cell(P, w[0], w[4] / w[0] * P.temp0));
    }
    __syncthreads();  // T is rewritten for the next plane only after every lane has read it
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = qn[m];
    tm = tc;
    tc = tn;
    e = en;
  }
  for (int off = 32; off > 0; off >>= 1) {
    vmin = dmin(vmin, __shfl_xor(vmin, off));
    tvmin = dmin(tvmin, __shfl_xor(tvmin, off));
  }
  if ((tid & 63) == 0) {
    atomic_min_pos(&st->vmin_bits, vmin);
    atomic_min_pos(&st->tvmin_bits, tvmin);
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_thermal_min(KParams P, long long n, const S *__restrict__ q, Status *st) {
  double v = 1e99;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    double qc[6], w[5];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(q[m * n + c]);
    const int e = q2w(P, qc, w);  // pressure(q) of timestep.f90:87
    if (e) report_next(st, e, (unsigned long long)c);
    v = dmin(v, tstep_cell(P, qc[0], w[4] / qc[0] * P.temp0));
  }
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) atomic_min_pos(&st->tvmin_bits, v);
}

__global__ void k_reset_tv(Status *st) {
  if (threadIdx.x == 0) st->tvmin_bits = (unsigned long long)__double_as_longlong(1e99);
}

CGeo make_cgeo(const evh_ctx *h) {
  CGeo g;
  g.nx = h->grid.nx;
  g.ny = h->grid.ny;
  g.nz = h->grid.nz;
  g.n = h->grid.n();
  g.zchunk = g.nz < 32 ? g.nz : 32;
  g.tiles_x = (g.nx + CX - 1) / CX;
  g.tiles_y = (g.ny + CY - 1) / CY;
  g.nchunks = (g.nz + g.zchunk - 1) / g.zchunk;
  g.ntiles = g.tiles_x * g.tiles_y * g.nchunks;
  return g;
}

}  // namespace

hipError_t launch_thermal_min(evh_ctx *h) {
  hipLaunchKernelGGL(k_reset_tv, dim3(1), dim3(64), 0, h->stream, h->st);
  const long long n = h->grid.n();
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (h->cfg.precision == EVH_FP32_MIXED)
    hipLaunchKernelGGL(k_thermal_min<float>, dim3((int)b), dim3(256), 0, h->stream, h->kp, n,
                       (const float *)h->qbuf[h->cur], h->st);
  else
    hipLaunchKernelGGL(k_thermal_min<double>, dim3((int)b), dim3(256), 0, h->stream, h->kp, n,
                       (const double *)h->qbuf[h->cur], h->st);
  return hipGetLastError();
}

// q[cur] -> q[1-cur]; the caller resets the CFL minimum before
hipError_t launch_conduction(evh_ctx *h, double dt) {
  hipLaunchKernelGGL(k_reset_tv, dim3(1), dim3(64), 0, h->stream, h->st);
  const CGeo g = make_cgeo(h);
  const int blocks = kXcds * ((g.ntiles + kXcds - 1) / kXcds);
  if (h->cfg.precision == EVH_FP32_MIXED)
    hipLaunchKernelGGL(k_conduct<float>, dim3(blocks), dim3(CT), 0, h->stream, h->kp, g,
                       (const float *)h->qbuf[h->cur], (float *)h->qbuf[1 - h->cur], dt, h->st);
  else
    hipLaunchKernelGGL(k_conduct<double>, dim3(blocks), dim3(CT), 0, h->stream, h->kp, g,
                       (const double *)h->qbuf[h->cur], (double *)h->qbuf[1 - h->cur], dt, h->st);
  return hipGetLastError();
}

}  // namespace evh
