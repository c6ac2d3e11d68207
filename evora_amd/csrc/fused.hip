// fused.hip -- the whole 3-D hydro step in one HBM pass.
//
// One kernel reads q^n once and writes q^{n+1} once; w, evo and s never
// exist in HBM.  A block owns a 16x12 column tile and marches up a chunk of
// z planes.  Waves 0-2 own the 192 interior columns one to one; wave 3 (the
// "ring wave") owns the 56 columns of the first halo ring and loads the
// second ring and the four corners, so every interior lane does useful work
// in every stage.  Per plane k:
//   S1  primitive (comp.f90:26-85) of plane k+1 for every owned column, with
//       the ghost cells of `bound` (boundary.f90) folded into the load index;
//       plane k goes to LDS with both halo rings.
//   S2  limited slopes in x, y, z and the predictor evo (muscl.f90:45-169)
//       of plane k; the x/y face states go to LDS per face; the interior
//       lanes solve their z face k-1/2 (riemann.f90) from registers.
//   S3  x and y faces of plane k: interior lane (i,j) solves its left and
//       bottom faces, the ring wave the 16 right-edge and 16 top-edge faces.
//       Each flux overwrites the two states it came from.
//   S4  plane k-1 is finished as q + dtx*(Fx) + dty*(Fy) + dtz*(Fz), in the
//       reference's x, y, z order (muscl.f90:191-272, integration :287), and
//       primitive + the CFL limit of the new state feed a min reduction for
//       the next hydrostep (timestep.f90:26-57).
// Template parameters: T compute type, S storage type of q (double/double
// for the FP64 form, float/float for the FP32-mixed form).
#include "evh_internal.h"
#include "physics_fast.h"

namespace evh {
namespace {

constexpr int TX = 16, TY = 12;        // interior tile
constexpr int NI = TX * TY;            // interior lanes (waves 0-2)
constexpr int NT = NI + 64;            // + the ring wave
constexpr int NR1 = 2 * TX + 2 * TY;   // first-ring columns (56)
constexpr int EW = TX + 4;             // extended plane pitch (x from -2 to TX+1)
constexpr int EA = EW * (TY + 4);      // extended plane area
constexpr int NFX = (TX + 1) * TY;     // x faces of a plane
constexpr int NFY = TX * (TY + 1);     // y faces of a plane

struct Geo {
  int nx, ny, nz;
  long long n;
  int zchunk;
  int tiles_x, tiles_y, nchunks;
  int ntiles;  // tiles_x * tiles_y * nchunks
};

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// Block b runs on XCD b % 8; giving XCD j the contiguous run of tiles
// [j*T/8, (j+1)*T/8) (x fastest, then y, then z chunk) keeps the halo
// columns a block shares with its x and y neighbours in one L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int j = b % kXcds, i = b / kXcds;
  const int base = ntiles / kXcds, rem = ntiles % kXcds;
  const int count = base + (j < rem ? 1 : 0);
  if (i >= count) return -1;
  return j * base + (j < rem ? j : rem) + i;
}

__device__ __forceinline__ int eidx(int x, int y) { return (x + 2) + EW * (y + 2); }

// mixed form: FP64 pressure recovery where the flow is fast
template <typename T>
struct Prim;

template <>
struct Prim<double> {
  __device__ static __forceinline__ int q2w(const KParams &P, double q[6], double w[5], double &rinv) {
    return fast::q2w(P, q, w, rinv);
  }
};

template <>
struct Prim<float> {
  // FP32 recovery, except where the subtraction E - kin cancels: there
  // (kin/E above the Mach threshold) p = g1 (E - kin) is evaluated in FP64.
  // With dualenergy on, the entropy branch (kin/E > de_thresh1,
  // comp.f90:61-64) recovers p = S rho**g1 and E = kin + p/g1 in FP64.
  __device__ static __forceinline__ int q2w(const KParams &P, float q[6], float w[5], float &rinv) {
    rinv = fm::rcp(q[0]);
    w[0] = q[0];
    w[1] = q[1] * rinv;
    w[2] = q[2] * rinv;
    w[3] = q[3] * rinv;
    const float m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const float kin = 0.5f * m2 * rinv;
    const bool fast_flow = kin > float(P.kin_frac64) * q[4];
    float pr;
    if (P.dualenergy && kin * fm::rcp(q[4]) > float(P.de_thresh1)) {
      const double r64 = double(q[0]);
      const double k64 = 0.5 * (double(q[1]) * q[1] + double(q[2]) * q[2] + double(q[3]) * q[3]) / r64;
      const double p64 = double(q[5]) * fm::pow_pos(r64, P.g1);
      q[4] = float(k64 + p64 * P.g1inv);
      pr = float(p64);
    } else {
      if (fast_flow) {
        const double k64 = 0.5 * (double(q[1]) * q[1] + double(q[2]) * q[2] + double(q[3]) * q[3]) / double(q[0]);
        pr = float(P.g1 * (double(q[4]) - k64));
      } else {
        pr = float(P.g1) * (q[4] - kin);
      }
      if (P.dualenergy) q[5] = pr * fm::pow_pos(q[0], -float(P.g1));
    }
    w[4] = pr;
    return (pr <= 0.f) ? kErrNegPressure : kOk;
  }
};

// load q at (gx, gy, gz) (ghost-mapped) and recover w
template <typename S, typename T>
__device__ __forceinline__ void load_cell(const KParams &P, const Geo &G, const S *__restrict__ q, int gx, int gy,
                                          int gz, T qc[6], T wc[5], T &rinv) {
  const bool per = P.boundtype == 2;
  const int cx = gx < -2 ? -2 : (gx > G.nx + 1 ? G.nx + 1 : gx);
  const int cy = gy < -2 ? -2 : (gy > G.ny + 1 ? G.ny + 1 : gy);
  const int sx = ghost_src(cx, G.nx, per);
  const int sy = ghost_src(cy, G.ny, per);
  const int sz = ghost_src(gz, G.nz, per);
  const long long c = (long long)sx + (long long)G.nx * ((long long)sy + (long long)G.ny * sz);
#pragma unroll
  for (int m = 0; m < 6; m++) qc[m] = T(q[m * G.n + c]);
  Prim<T>::q2w(P, qc, wc, rinv);
  if (P.boundtype == 3 && cx >= G.nx) {  // reflecting upper x (boundary.f90:57-64)
    wc[1] = -wc[1];
    wc[2] = -wc[2];
    wc[3] = -wc[3];
  }
}

template <typename T>
__device__ __forceinline__ void store_w(T (*W)[EA], int slot, const T w[5]) {
#pragma unroll
  for (int m = 0; m < 5; m++) W[m][slot] = w[m];
}

template <typename T, typename S, int SOLVER, int LIMITER>
__global__ void __launch_bounds__(NT) k_fused(KParams P, Geo G, const S *__restrict__ qin, S *__restrict__ qout,
                                              double dt, Status *st) {
  __shared__ T W[5][EA];      // w of plane k, both halo rings
  __shared__ T XB[10][NFX];   // x face (fi, j), fi = 0..TX: [0,5) left state, [5,10) right state; then flux
  __shared__ T YB[10][NFY];   // y face (i, fj), fj = 0..TY

  const int tid = threadIdx.x;
  const bool ring = tid >= NI;
  // tile of this block
  const int lin = xcd_tile(blockIdx.x, G.ntiles);
  if (lin < 0) return;  // padding block (whole block exits before any barrier)
  const int tile = lin % (G.tiles_x * G.tiles_y);
  const int chunk = lin / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * TX, y0 = (tile / G.tiles_x) * TY;
  const int z0 = chunk * G.zchunk;
  const int z1 = min(z0 + G.zchunk, G.nz);

  // owned column (interior cell, or first-ring cell for the ring wave); the
  // ring wave also loads one second-ring cell or first-ring corner per lane
  int cx, cy, side = -1, r2x = 0, r2y = 0;
  const int lane = tid - NI;
  if (!ring) {
    cx = tid % TX;
    cy = tid / TX;
  } else if (lane < TY) {
    side = 0; cx = -1; cy = lane; r2x = -2; r2y = lane;
  } else if (lane < 2 * TY) {
    side = 1; cx = TX; cy = lane - TY; r2x = TX + 1; r2y = cy;
  } else if (lane < 2 * TY + TX) {
    side = 2; cx = lane - 2 * TY; cy = -1; r2x = cx; r2y = -2;
  } else if (lane < NR1) {
    side = 3; cx = lane - 2 * TY - TX; cy = TY; r2x = cx; r2y = TY + 1;
  } else {  // lanes NR1..NR1+3 load the first-ring corners, the rest idle
    const int c = lane - NR1;
    cx = (c & 1) ? TX : -1;
    cy = (c & 2) ? TY : -1;
    r2x = cx;
    r2y = cy;
    if (c >= 4) r2x = -3;  // no second-ring job
  }
  const bool owns_col = !ring || side >= 0;
  const int gx = x0 + cx, gy = y0 + cy;
  const bool inside = !ring && gx < G.nx && gy < G.ny;
  const int me = eidx(cx, cy);

  const T dt2 = T(-0.5 * dt);
  const T dtx = T(dt * P.dxinv), dty = T(dt * P.dyinv), dtz = T(dt * P.dzinv);
  const T half = T(0.5);

  // FP32 keeps the base q of the update in registers; FP64 reloads it from
  // cache in S4 (it was read one plane earlier), which saves 24 VGPRs
  constexpr bool kKeepQ = sizeof(T) == 4;
  T wm[5], wc[5], wp[5], qn[6], qc[6], tq[6];
  T rm, rc, rp;
  T part[6], zr_prev[5], fz_prev[6];
  double vmin = 1e6;

  load_cell<S, T>(P, G, qin, gx, gy, z0 - 2, tq, wm, rm);
  load_cell<S, T>(P, G, qin, gx, gy, z0 - 1, tq, wc, rc);
#pragma unroll
  for (int m = 0; m < 6; m++) { part[m] = T(0); fz_prev[m] = T(0); qc[m] = T(0); qn[m] = T(0); }
#pragma unroll
  for (int m = 0; m < 5; m++) zr_prev[m] = T(0);

  for (int k = z0 - 1; k <= z1; k++) {
    const bool out_plane = k >= z0 && k < z1;
    // ---- S1: primitive of plane k+1 (own column); rings of plane k to LDS
    if (owns_col) {
      if (kKeepQ) load_cell<S, T>(P, G, qin, gx, gy, k + 1, qn, wp, rp);
      else load_cell<S, T>(P, G, qin, gx, gy, k + 1, tq, wp, rp);
    }
    if (ring && r2x > -3) {
      T rw[5], rr;
      load_cell<S, T>(P, G, qin, x0 + r2x, y0 + r2y, k, tq, rw, rr);
      store_w(W, eidx(r2x, r2y), rw);
    }
    if (owns_col) store_w(W, me, wc);
    __syncthreads();

    // ---- S2: slopes, predictor and face states of plane k
    T sx[5], sy[5], sz[5], ev[5];
    T fz[6];
    if (owns_col) {
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sx[m] = fast::slope3<LIMITER>(W[m][me - 1], wc[m], W[m][me + 1]);
      sy[m] = fast::slope3<LIMITER>(W[m][me - EW], wc[m], W[m][me + EW]);
      sz[m] = fast::slope3<LIMITER>(wm[m], wc[m], wp[m]);
    }
    if (P.secondorder) {
      T a1[5], a2[5], a3[5];
      fast::predictor_dir<0>(P, wc, rc, sx, a1);
      fast::predictor_dir<1>(P, wc, rc, sy, a2);
      fast::predictor_dir<2>(P, wc, rc, sz, a3);
      const T ix = T(P.dxinv), iy = T(P.dyinv), iz = T(P.dzinv);
#pragma unroll
      for (int m = 0; m < 5; m++) ev[m] = dt2 * (a1[m] * ix + a2[m] * iy + a3[m] * iz);
    } else {
#pragma unroll
      for (int m = 0; m < 5; m++) { ev[m] = T(0); sx[m] = T(0); sy[m] = T(0); sz[m] = T(0); }
    }
    // face states (muscl.f90:207-210): w + 0.5 s + evo on the high side, w - 0.5 s + evo on the low side
    if (!ring) {
      const int fxl = cx + (TX + 1) * cy, fyb = cx + TX * cy;
#pragma unroll
      for (int m = 0; m < 5; m++) {
        XB[5 + m][fxl] = wc[m] - half * sx[m] + ev[m];       // right state of the left face
        XB[m][fxl + 1] = wc[m] + half * sx[m] + ev[m];       // left state of the right face
        YB[5 + m][fyb] = wc[m] - half * sy[m] + ev[m];
        YB[m][fyb + TX] = wc[m] + half * sy[m] + ev[m];
      }
    } else if (side == 0) {
#pragma unroll
      for (int m = 0; m < 5; m++) XB[m][(TX + 1) * cy] = wc[m] + half * sx[m] + ev[m];
    } else if (side == 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) XB[5 + m][TX + (TX + 1) * cy] = wc[m] - half * sx[m] + ev[m];
    } else if (side == 2) {
#pragma unroll
      for (int m = 0; m < 5; m++) YB[m][cx] = wc[m] + half * sy[m] + ev[m];
    } else {
#pragma unroll
      for (int m = 0; m < 5; m++) YB[5 + m][cx + TX * TY] = wc[m] - half * sy[m] + ev[m];
    }
    // z face k-1/2 (interior lanes, from registers)
    if (!ring) {
      T zl[5], zr[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        zl[m] = wc[m] - half * sz[m] + ev[m];
        zr[m] = wc[m] + half * sz[m] + ev[m];
      }
      if (k >= z0) {
        const int e = fast::riemann<SOLVER, 3>(P, zr_prev, zl, fz);
        if (e && inside) report(st, e, (unsigned long long)(gx + (long long)G.nx * (gy + (long long)G.ny * k)));
      }
#pragma unroll
      for (int m = 0; m < 5; m++) zr_prev[m] = zr[m];
    }
    }  // owns_col
    __syncthreads();

    // ---- S3: x and y faces of plane k
    if (out_plane) {
      int fxi = -1, fyi = -1;
      if (!ring) {
        fxi = cx + (TX + 1) * cy;
        fyi = cx + TX * cy;
      } else {
        if (lane < TY) fxi = TX + (TX + 1) * lane;          // right-edge x faces
        else if (lane < TY + TX) fyi = (lane - TY) + TX * TY;  // top-edge y faces
      }
      if (fxi >= 0) {
        T wl[5], wr[5], f[6];
#pragma unroll
        for (int m = 0; m < 5; m++) { wl[m] = XB[m][fxi]; wr[m] = XB[5 + m][fxi]; }
        const int e = fast::riemann<SOLVER, 1>(P, wl, wr, f);
        const int fgx = x0 + fxi % (TX + 1), fgy = y0 + fxi / (TX + 1);
        if (e && fgx <= G.nx && fgy < G.ny) report(st, e, (unsigned long long)(min(fgx, G.nx - 1) + (long long)G.nx * (fgy + (long long)G.ny * k)));
#pragma unroll
        for (int m = 0; m < 6; m++) XB[m][fxi] = f[m];
      }
      if (fyi >= 0) {
        T wl[5], wr[5], f[6];
#pragma unroll
        for (int m = 0; m < 5; m++) { wl[m] = YB[m][fyi]; wr[m] = YB[5 + m][fyi]; }
        const int e = fast::riemann<SOLVER, 2>(P, wl, wr, f);
        const int fgx = x0 + fyi % TX, fgy = y0 + fyi / TX;
        if (e && fgx < G.nx && fgy <= G.ny) report(st, e, (unsigned long long)(fgx + (long long)G.nx * (min(fgy, G.ny - 1) + (long long)G.ny * k)));
#pragma unroll
        for (int m = 0; m < 6; m++) YB[m][fyi] = f[m];
      }
    }
    __syncthreads();

    // ---- S4: finish plane k-1, x/y differences of plane k
    if (!ring) {
      if (k >= z0) {
        if (k - 1 >= z0 && inside) {
          T o[6];
#pragma unroll
          for (int m = 0; m < 6; m++) o[m] = part[m] + dtz * (fz_prev[m] - fz[m]);
          const long long c = (long long)gx + (long long)G.nx * ((long long)gy + (long long)G.ny * (k - 1));
#pragma unroll
          for (int m = 0; m < 6; m++) qout[m * G.n + c] = S(o[m]);
          T on[6], wn[5], rn;
#pragma unroll
          for (int m = 0; m < 6; m++) on[m] = T(S(o[m]));
          const int ep = Prim<T>::q2w(P, on, wn, rn);
          if (ep) report_next(st, ep, (unsigned long long)c);
          vmin = dmin(vmin, fast::cfl_cell(P, wn, rn));
        }
#pragma unroll
        for (int m = 0; m < 6; m++) fz_prev[m] = fz[m];
      }
      if (out_plane) {
        const int fxl = cx + (TX + 1) * cy, fyb = cx + TX * cy;
        if (!kKeepQ) {  // base q(k) of this column, dual-energy corrected as in S1
          T wt[5], rt;
          load_cell<S, T>(P, G, qin, gx, gy, k, qc, wt, rt);
        }
#pragma unroll
        for (int m = 0; m < 6; m++) part[m] = qc[m] + dtx * (XB[m][fxl] - XB[m][fxl + 1]);
#pragma unroll
        for (int m = 0; m < 6; m++) part[m] = part[m] + dty * (YB[m][fyb] - YB[m][fyb + TX]);
      }
    }
    // rotate the column window
#pragma unroll
    for (int m = 0; m < 5; m++) { wm[m] = wc[m]; wc[m] = wp[m]; }
    rm = rc;
    rc = rp;
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = qn[m];
  }

  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((tid & 63) == 0) atomic_min_pos(&st->vmin_bits, vmin);
}

// primitive + CFL min over q without materialising w (first step, after
// uploads): exact IEEE arithmetic, so the first dt equals the reference's.
template <typename S>
__global__ void __launch_bounds__(256) k_cfl(KParams P, Geo G, const S *__restrict__ q, Status *st) {
  double vmin = 1e6;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < G.n; c += (long long)gridDim.x * 256) {
    double qc[6], wc[5];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(q[m * G.n + c]);
    const int e = q2w(P, qc, wc);
    if (e) report_next(st, e, (unsigned long long)c);
    vmin = dmin(vmin, cfl_cell(P, wc));
  }
  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((threadIdx.x & 63) == 0) atomic_min_pos(&st->vmin_bits, vmin);
}

Geo make_geo(const evh_ctx *h) {
  Geo g;
  g.nx = h->grid.nx;
  g.ny = h->grid.ny;
  g.nz = h->grid.nz;
  g.n = h->grid.n();
  g.zchunk = h->zchunk > 0 ? h->zchunk : 128;
  if (g.zchunk > g.nz) g.zchunk = g.nz;
  g.tiles_x = (g.nx + TX - 1) / TX;
  g.tiles_y = (g.ny + TY - 1) / TY;
  g.nchunks = (g.nz + g.zchunk - 1) / g.zchunk;
  g.ntiles = g.tiles_x * g.tiles_y * g.nchunks;
  return g;
}

template <typename T, typename S, int SOLVER, int LIMITER>
void launch_fused_t(evh_ctx *h, double dt) {
  const Geo g = make_geo(h);
  const int blocks = kXcds * ((g.ntiles + kXcds - 1) / kXcds);
  hipLaunchKernelGGL((k_fused<T, S, SOLVER, LIMITER>), dim3(blocks), dim3(NT), 0, h->stream, h->kp, g,
                     (const S *)h->qbuf[h->cur], (S *)h->qbuf[1 - h->cur], dt, h->st);
}

template <typename T, typename S>
hipError_t launch_fused_prec(evh_ctx *h, double dt) {
  const int s = h->kp.solver, l = h->kp.limiter;
  if (s == 1 && l == 1) launch_fused_t<T, S, 1, 1>(h, dt);
  else if (s == 1 && l == 2) launch_fused_t<T, S, 1, 2>(h, dt);
  else if (s == 1 && l == 3) launch_fused_t<T, S, 1, 3>(h, dt);
  else if (s == 2 && l == 1) launch_fused_t<T, S, 2, 1>(h, dt);
  else if (s == 2 && l == 2) launch_fused_t<T, S, 2, 2>(h, dt);
  else if (s == 2 && l == 3) launch_fused_t<T, S, 2, 3>(h, dt);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace

bool fused_supported(const evh_config &c) { return c.nx >= 2 && c.ny >= 2 && c.nz >= 2; }

hipError_t launch_cfl(evh_ctx *h) {
  const Geo g = make_geo(h);
  long long b = (g.n + 255) / 256;
  if (b > 8192) b = 8192;
  if (h->cfg.precision == EVH_FP32_MIXED)
    hipLaunchKernelGGL(k_cfl<float>, dim3((int)b), dim3(256), 0, h->stream, h->kp, g, (const float *)h->qbuf[h->cur],
                       h->st);
  else
    hipLaunchKernelGGL(k_cfl<double>, dim3((int)b), dim3(256), 0, h->stream, h->kp, g,
                       (const double *)h->qbuf[h->cur], h->st);
  return hipGetLastError();
}

hipError_t launch_fused_step(evh_ctx *h, double dt) {
  if (h->cfg.precision == EVH_FP32_MIXED) return launch_fused_prec<float, float>(h, dt);
  return launch_fused_prec<double, double>(h, dt);
}

}  // namespace evh
