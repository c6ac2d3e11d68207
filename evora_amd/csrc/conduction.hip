// conduction.hip -- thermal conduction (conduction.f90:26-164) and its time
// step limit (timestep.f90:72-104) on the device.
//
// The temperature T = p / rho * temp0 (comp.f90:96-100, no chemistry) of
// every cell is stored by the fused sweep that produced the state (or by
// k_temp after an upload).  k_conduct is then one HBM pass over q in place:
// a block owns a 64 x 8 column tile and marches a chunk of z planes; per
// plane it stages the T of the tile and of its one-cell ring in LDS (the
// ghost cells of `bound` folded into the load index; z neighbours stay in
// registers), adds the x, y and z temperature-flux differences of each
// interior cell to q(5) and q(6) (thermalintegration), writes back those
// two components only, and feeds the CFL limit of `hydrostep` and the
// conduction limit of `thermalstep` of the new state into the status block,
// so the next step needs no pass of its own.  The update never reads another
// cell's q, only its T, so the pass runs in place.  Only simplekappa:
// kappa_eff = 1/g1.  k_thermal_min gives the conduction limit of a state that
// did not come out of k_conduct (after an upload or a sweep).
//
// Built twice.  EVH_COND_EXACT=1 (compiled without FMA contraction) is the
// reference's arithmetic: IEEE division, the direction-by-direction update
// of thermalintegration, and the powers T**2.5 and q(1)**g1 from the device
// pow (which may differ from the host's in the last bit).  The default build
// is the form of the fast fused step: one refined reciprocal of rho per
// cell, the three directions' energy changes summed before they enter q(5)
// and q(6), q(1)**(-g1) from the cube-root power, T**2.5 as T*T*sqrt(T), and
// the single-reciprocal CFL limit (results within a few ulp).
#include "evh_internal.h"
#include "physics_fast.h"

#ifndef EVH_COND_EXACT
#define EVH_COND_EXACT 0
#endif

namespace evh {
namespace {

constexpr bool kExact = EVH_COND_EXACT != 0;
#ifndef EVH_COND_CX
#define EVH_COND_CX 64
#endif
#ifndef EVH_COND_CY
#define EVH_COND_CY 8
#endif
#ifndef EVH_COND_ZCHUNK
#define EVH_COND_ZCHUNK 128
#endif
constexpr int CX = EVH_COND_CX, CY = EVH_COND_CY, CT = CX * CY;  // tile and threads
constexpr int RX = CX + 2, RY = CY + 2;        // with the one-cell ring
constexpr int NRING = 2 * CX + 2 * CY;         // ring cells (no corners)

struct CGeo {
  int nx, ny, nz;
  long long n;
  int zchunk, tiles_x, tiles_y, nchunks, ntiles;
  // a z-slab (evh_slab): T of the neighbours' planes next to the slab, below
  // (tlo) and above (thi), [ny][nx] in the storage type; null: bound's rules
  const void *tlo, *thi;
};

// linear index of the cell (gx, gy, gz) or of the interior cell that
// `bound` (boundary.f90) copies into it: mirror or periodic source; the
// upper-x reflection of boundtype 3 negates velocities only, which T does not
// see.  Indices beyond the ghost layer (partial tiles) are clamped: those
// lanes update nothing.
__device__ __forceinline__ long long src_cell(const CGeo &G, bool per, int gx, int gy, int gz) {
  gx = min(max(gx, -2), G.nx + 1);
  gy = min(max(gy, -2), G.ny + 1);
  gz = min(max(gz, -2), G.nz + 1);
  // an axis of one cell (1-D and 2-D grids) has that cell as every ghost:
  // the two-layer rules would name cells -1 or 1 of it
  const int sx = G.nx == 1 ? 0 : ghost_src(gx, G.nx, per);
  const int sy = G.ny == 1 ? 0 : ghost_src(gy, G.ny, per);
  const int sz = G.nz == 1 ? 0 : ghost_src(gz, G.nz, per);
  return (long long)sx + (long long)G.nx * ((long long)sy + (long long)G.ny * sz);
}

// T of the cell (gx, gy) of plane gz: from tb, or in a z-slab, for the plane
// beyond a side with a neighbour, from that neighbour's plane (the one
// communicate passes before thermalconduction, main.f90:173)
template <typename S>
__device__ __forceinline__ double t_at(const CGeo &G, bool per, const S *__restrict__ tb, int gx, int gy, int gz) {
  if (gz < 0 && G.tlo) return double(((const S *)G.tlo)[src_cell(G, per, gx, gy, 0)]);
  if (gz >= G.nz && G.thi) return double(((const S *)G.thi)[src_cell(G, per, gx, gy, 0)]);
  return double(tb[src_cell(G, per, gx, gy, gz)]);
}

// Per-cell arithmetic of the pass.  temp: T of a cell from its conserved q
// (pressure, comp.f90:45-85; q is not rewritten: conduction refuses dual
// energy).  update: the three directions of conduction.f90:26-123 for one
// cell from its temperature and its six neighbours'.  limits: the next
// step's hydrostep and thermalstep minima from the new q of a cell.
template <bool EXACT>
struct CondMath;

template <>
struct CondMath<true> {
  struct Coef {
    double dx_delta, dy_delta, dz_delta;
  };
  __device__ static Coef coef(const KParams &P, const CGeo &, double dt) {
    return {dt * P.dxinv, dt * P.dyinv, dt * P.dzinv};
  }
  __device__ static __forceinline__ double temp(const KParams &P, const double qin[6], int &err) {
    double q[6], w[5];
#pragma unroll
    for (int m = 0; m < 6; m++) q[m] = qin[m];
    err = q2w(P, q, w);
    return w[4] / w[0] * P.temp0;
  }
  // thermalflux with simplekappa (conduction.f90:126-144)
  __device__ static __forceinline__ double tflux(const KParams &P, double tl, double tr, double dinv) {
    const double kappa_eff = 1.0 / P.g1;
    const double dtempdx = (tr - tl) * dinv;
    return -kappa_eff * dtempdx;
  }
  // thermalintegration (conduction.f90:146-162); pg = q(1)**g1
  __device__ static __forceinline__ void tintegrate(const KParams &P, double jl, double jr, double q[6],
                                                    double delta, double pg) {
    const double de = delta * (jl - jr);
    q[4] = q[4] + de;
    q[5] = q[5] + P.g1 * de / pg;
  }
  // returns the pressure check of the primitive before the pass (main.f90:171)
  __device__ static __forceinline__ int update(const KParams &P, const CGeo &G, const Coef &k, double q[6],
                                               double tc, double txl, double txr, double tyl, double tyr,
                                               double tzl, double tzr) {
    int e;
    temp(P, q, e);
    const double pg = pow(q[0], P.g1);
    double jl = tflux(P, txl, tc, P.dxinv), jr = tflux(P, tc, txr, P.dxinv);
    if (G.nx != 1) tintegrate(P, jl, jr, q, k.dx_delta, pg);
    jl = tflux(P, tyl, tc, P.dyinv);
    jr = tflux(P, tc, tyr, P.dyinv);
    if (G.ny != 1) tintegrate(P, jl, jr, q, k.dy_delta, pg);
    jl = tflux(P, tzl, tc, P.dzinv);
    jr = tflux(P, tc, tzr, P.dzinv);
    if (G.nz != 1) tintegrate(P, jl, jr, q, k.dz_delta, pg);
    return e;
  }
  // min over cells of rho / kappa, kappa = kappa0 T**2.5 (timestep.f90:84-93)
  __device__ static __forceinline__ double tstep_cell(const KParams &P, double rho, double temp) {
    return rho / (P.kappa0 * pow(temp, 2.5));
  }
  __device__ static __forceinline__ int limits(const KParams &P, double q[6], double &v, double &tv) {
    double w[5];
    const int e = q2w(P, q, w);
    v = cfl_cell(P, w);
    tv = tstep_cell(P, w[0], w[4] / w[0] * P.temp0);
    return e;
  }
};

template <>
struct CondMath<false> {
  // de = sum over directions of dt dinv^2 / g1 ((tr - tc) - (tc - tl)), the
  // reference's delta * (jl - jr) with jl, jr = -(1/g1) dT/dx
  struct Coef {
    double cx, cy, cz;
  };
  __device__ static Coef coef(const KParams &P, const CGeo &G, double dt) {
    const double kd = dt / P.g1;
    return {G.nx != 1 ? kd * P.dxinv * P.dxinv : 0.0, G.ny != 1 ? kd * P.dyinv * P.dyinv : 0.0,
            G.nz != 1 ? kd * P.dzinv * P.dzinv : 0.0};
  }
  __device__ static __forceinline__ double temp(const KParams &P, const double q[6], int &err) {
    const double rinv = fm::rcp(q[0]);
    const double m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const double pr = P.g1 * (q[4] - 0.5 * m2 * rinv);
    err = (pr <= 0.0) ? kErrNegPressure : kOk;
    return pr * rinv * P.temp0;
  }
  __device__ static __forceinline__ int update(const KParams &P, const CGeo &, const Coef &k, double q[6],
                                               double tc, double txl, double txr, double tyl, double tyr,
                                               double tzl, double tzr) {
    const double m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const int e = (P.g1 * (q[4] - 0.5 * m2 * fm::rcp(q[0])) <= 0.0) ? kErrNegPressure : kOk;
    const double de = k.cx * ((txr - tc) - (tc - txl)) + k.cy * ((tyr - tc) - (tc - tyl)) +
                      k.cz * ((tzr - tc) - (tc - tzl));
    q[4] = q[4] + de;
    q[5] = q[5] + P.g1 * de * fast::rpow_g1(P, q[0]);
    return e;
  }
  __device__ static __forceinline__ int limits(const KParams &P, double q[6], double &v, double &tv) {
    double w[5], rinv;
    const int e = fast::q2w(P, q, w, rinv);
    v = fast::cfl_cell(P, w, rinv);
    const double t = w[4] * rinv * P.temp0;
    tv = w[0] * fm::rcp(P.kappa0 * (t * t) * fm::sqrt_pos(t));
    return e;
  }
};

using M = CondMath<kExact>;

template <typename S>
__global__ void __launch_bounds__(CT) k_conduct(KParams P, CGeo G, S *__restrict__ q, const S *__restrict__ tb,
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
  const bool per = P.boundtype == 2;
  // ring job: lanes 0..NRING-1 each own one ring cell
  int rx = 0, ry = 0;
  const bool ring = tid < NRING;
  if (tid < CX) { rx = tid; ry = -1; }
  else if (tid < 2 * CX) { rx = tid - CX; ry = CY; }
  else if (tid < 2 * CX + CY) { rx = -1; ry = tid - 2 * CX; }
  else if (tid < NRING) { rx = CX; ry = tid - 2 * CX - CY; }

  const typename M::Coef kc = M::coef(P, G, dt);
  double vmin = 1e6, tvmin = 1e99;
  const long long col = (long long)min(gx, G.nx - 1) + (long long)G.nx * min(gy, G.ny - 1);
  const long long plane = (long long)G.nx * G.ny;

  // tm, tc, tn: T of planes k-1, k, k+1 of the column; qn: q of plane k,
  // loaded one plane ahead
  double tm = t_at(G, per, tb, gx, gy, z0 - 1);
  double tc = double(tb[src_cell(G, per, gx, gy, z0)]);
  S qn[6];
#pragma unroll
  for (int m = 0; m < 6; m++) qn[m] = q[m * G.n + col + plane * z0];
  for (int k = z0; k < z1; k++) {
    double qc[6];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(qn[m]);
    if (k + 1 < z1) {
#pragma unroll
      for (int m = 0; m < 6; m++) qn[m] = q[m * G.n + col + plane * (k + 1)];
    }
    const double tn = t_at(G, per, tb, gx, gy, k + 1);
    T[(cx + 1) + RX * (cy + 1)] = tc;
    if (ring) T[(rx + 1) + RX * (ry + 1)] = double(tb[src_cell(G, per, x0 + rx, y0 + ry, k)]);
    __syncthreads();
    if (inside) {
      const int me = (cx + 1) + RX * (cy + 1);
      const long long c = col + plane * k;
      // the primitive of main.f90:171 before thermalconduction stops on a
      // non-positive pressure of this cell
      const int e = M::update(P, G, kc, qc, tc, T[me - 1], T[me + 1], T[me - RX], T[me + RX], tm, tn);
      if (e) report(st, e, (unsigned long long)c);
      q[4 * G.n + c] = S(qc[4]);
      q[5 * G.n + c] = S(qc[5]);
      // the next step's limits from the new state (as stored)
      double qs[6], v, tv;
#pragma unroll
      for (int m = 0; m < 6; m++) qs[m] = double(S(qc[m]));
      const int ep = M::limits(P, qs, v, tv);
      if (ep) report_next(st, ep, (unsigned long long)c);
      vmin = dmin(vmin, v);
      tvmin = dmin(tvmin, tv);
    }
    __syncthreads();  // T is rewritten for the next plane only after every lane has read it
    tm = tc;
    tc = tn;
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

// T of every cell of q (the state did not come out of a fused sweep)
template <typename S>
__global__ void __launch_bounds__(256) k_temp(KParams P, long long n, const S *__restrict__ q, S *__restrict__ tb) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    double qc[6];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(q[m * n + c]);
    int e;
    tb[c] = S(M::temp(P, qc, e));
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_thermal_min(KParams P, long long n, const S *__restrict__ q, Status *st) {
  double v = 1e99;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    double qc[6], vc, tv;
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(q[m * n + c]);
    const int e = M::limits(P, qc, vc, tv);  // pressure(q) of timestep.f90:87
    if (e) report_next(st, e, (unsigned long long)c);
    v = dmin(v, tv);
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
  g.tiles_x = (g.nx + CX - 1) / CX;
  g.tiles_y = (g.ny + CY - 1) / CY;
  g.zchunk = z_chunk((long long)g.tiles_x * g.tiles_y, g.nz, EVH_COND_ZCHUNK);
  if (g.zchunk > g.nz) g.zchunk = g.nz;
  g.nchunks = (g.nz + g.zchunk - 1) / g.zchunk;
  g.ntiles = g.tiles_x * g.tiles_y * g.nchunks;
  g.tlo = h->zslab[0] ? h->thalo[0] : nullptr;
  g.thi = h->zslab[1] ? h->thalo[1] : nullptr;
  return g;
}

template <typename S>
void launch_temp_s(evh_ctx *h) {
  const long long n = h->grid.n();
  const long long b = std::min((n + 255) / 256, 8192LL);
  hipLaunchKernelGGL(k_temp<S>, dim3((int)b), dim3(256), 0, h->stream, h->kp, n, (const S *)h->qbuf[h->cur],
                     (S *)h->tbuf);
}

template <typename S>
void launch_thermal_min_s(evh_ctx *h) {
  hipLaunchKernelGGL(k_reset_tv, dim3(1), dim3(64), 0, h->stream, h->st);
  const long long n = h->grid.n();
  const long long b = std::min((n + 255) / 256, 8192LL);
  hipLaunchKernelGGL(k_thermal_min<S>, dim3((int)b), dim3(256), 0, h->stream, h->kp, n, (const S *)h->qbuf[h->cur],
                     h->st);
}

// q[cur] in place, with T of q[cur] in tbuf; the caller resets the CFL
// minimum before
template <typename S>
void launch_conduction_s(evh_ctx *h, double dt) {
  hipLaunchKernelGGL(k_reset_tv, dim3(1), dim3(64), 0, h->stream, h->st);
  const CGeo g = make_cgeo(h);
  const int blocks = xcd_blocks(g.ntiles);
  hipLaunchKernelGGL(k_conduct<S>, dim3(blocks), dim3(CT), 0, h->stream, h->kp, g, (S *)h->qbuf[h->cur],
                     (const S *)h->tbuf, dt, h->st);
}

}  // namespace

#if EVH_COND_EXACT
// the reference's arithmetic (FP64 only: the exact and staged modes)
hipError_t launch_thermal_min_exact(evh_ctx *h) {
  launch_thermal_min_s<double>(h);
  return hipGetLastError();
}
hipError_t launch_temp_exact(evh_ctx *h) {
  launch_temp_s<double>(h);
  return hipGetLastError();
}
hipError_t launch_conduction_exact(evh_ctx *h, double dt) {
  launch_conduction_s<double>(h, dt);
  return hipGetLastError();
}
#else
// the fast form in EVH_MODE_FUSED; the reference's arithmetic otherwise
static bool use_exact(const evh_ctx *h) { return h->mode != EVH_MODE_FUSED && h->cfg.precision == EVH_FP64; }

hipError_t launch_thermal_min(evh_ctx *h) {
  if (use_exact(h)) return launch_thermal_min_exact(h);
  if (h->cfg.precision == EVH_FP32_MIXED) launch_thermal_min_s<float>(h);
  else launch_thermal_min_s<double>(h);
  return hipGetLastError();
}
hipError_t launch_temp(evh_ctx *h) {
  if (use_exact(h)) return launch_temp_exact(h);
  if (h->cfg.precision == EVH_FP32_MIXED) launch_temp_s<float>(h);
  else launch_temp_s<double>(h);
  return hipGetLastError();
}
hipError_t launch_conduction(evh_ctx *h, double dt) {
  if (use_exact(h)) return launch_conduction_exact(h, dt);
  if (h->cfg.precision == EVH_FP32_MIXED) launch_conduction_s<float>(h, dt);
  else launch_conduction_s<double>(h, dt);
  return hipGetLastError();
}
#endif

}  // namespace evh
