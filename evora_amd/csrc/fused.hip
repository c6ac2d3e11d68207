// fused.hip -- the whole 3-D hydro step in one HBM pass.
//
// One kernel reads q^n once and writes q^{n+1} once.  Each block owns a
// 14x14 column tile (plus a one-cell halo ring, 16x16 threads = one column
// each) and marches up a chunk of z planes, keeping the primitive state of
// planes k-1, k, k+1 of its column in registers and plane k with a two-cell
// ring in LDS.  Per plane it
//   1. recovers w = q2w(q) of plane k+1 (primitive, comp.f90:26-85) with the
//      ghost cells of `bound` folded into the load index (boundary.f90),
//   2. forms the limited slopes in x, y, z and the half-step predictor evo
//      (compevo, muscl.f90:45-113) for plane k,
//   3. solves the x and y faces of plane k and the z face k-1/2
//      (riemann.f90), and
//   4. finishes plane k-1: q + dtx*(Fx) then + dty*(Fy) then + dtz*(Fz) in
//      the reference's x, y, z order (muscl.f90:191-272, integration :287),
//      and folds primitive + the CFL limit of the new state into a min
//      reduction for the next step's hydrostep (timestep.f90:26-57).
// w, evo and s never exist in HBM.
#include "evh_internal.h"

namespace evh {
namespace {

constexpr int BX = 16, BY = 16;       // threads per block (columns incl. halo ring)
constexpr int TX = BX - 2, TY = BY - 2;  // interior tile
constexpr int NT = BX * BY;
constexpr int NRING = 2 * TX + 2 * TY;   // second halo ring (x/y slopes of the first ring)
constexpr int PW = BX + 2;               // LDS plane pitch (ring at -1 and BX)

struct Geo {
  int nx, ny, nz;
  long long n;
  int zchunk;
  int tiles_x;
};

// load q at (gx, gy, gz) (ghost-mapped) and recover w; returns the q used (dual-energy fixed)
template <typename S, typename T>
__device__ __forceinline__ void load_cell(const KParams &P, const Geo &G, const S *__restrict__ q, int gx, int gy,
                                          int gz, T qc[6], T wc[5]) {
  const bool per = P.boundtype == 2;
  const int cx = gx < -2 ? -2 : (gx > G.nx + 1 ? G.nx + 1 : gx);
  const int cy = gy < -2 ? -2 : (gy > G.ny + 1 ? G.ny + 1 : gy);
  const int sx = ghost_src(cx, G.nx, per);
  const int sy = ghost_src(cy, G.ny, per);
  const int sz = ghost_src(gz, G.nz, per);
  const long long c = (long long)sx + (long long)G.nx * ((long long)sy + (long long)G.ny * sz);
#pragma unroll
  for (int m = 0; m < 6; m++) qc[m] = T(q[m * G.n + c]);
  q2w(P, qc, wc);
  if (P.boundtype == 3 && cx >= G.nx) {  // reflecting upper x (boundary.f90:57-64)
    wc[1] = -wc[1];
    wc[2] = -wc[2];
    wc[3] = -wc[3];
  }
}

template <int SOLVER, int LIMITER, typename S>
__global__ void __launch_bounds__(NT) k_fused(KParams P, Geo G, const S *__restrict__ qin, S *__restrict__ qout,
                                              double dt, Status *st) {
  using T = double;
  __shared__ T wpl[5][PW * PW];   // plane k, with the second ring
  __shared__ T xs[5][NT];         // right-face x states
  __shared__ T ys[5][NT];         // top-face y states
  __shared__ T fx[6][NT];         // flux through each column's left x face
  __shared__ T fy[6][NT];         // flux through each column's bottom y face

  const int tid = threadIdx.x;
  const int tx = tid % BX, ty = tid / BX;
  const int bx = blockIdx.x % G.tiles_x, by = blockIdx.x / G.tiles_x;
  const int x0 = bx * TX, y0 = by * TY;
  const int gx = x0 - 1 + tx, gy = y0 - 1 + ty;
  const int z0 = blockIdx.y * G.zchunk;
  const int z1 = min(z0 + G.zchunk, G.nz);
  const bool interior = tx >= 1 && tx <= TX && ty >= 1 && ty <= TY && gx < G.nx && gy < G.ny;

  // ring cell of this thread (tid < NRING)
  int rgx = 0, rgy = 0, rslot = -1;
  if (tid < NRING) {
    int r = tid;
    if (r < TY) { rgx = x0 - 2; rgy = y0 + r; rslot = (r + 2) * PW + 0; }
    else if ((r -= TY) < TY) { rgx = x0 + TX + 1; rgy = y0 + r; rslot = (r + 2) * PW + (BX + 1); }
    else if ((r -= TY) < TX) { rgx = x0 + r; rgy = y0 - 2; rslot = 0 * PW + (r + 2); }
    else { r -= TX; rgx = x0 + r; rgy = y0 + TY + 1; rslot = (BY + 1) * PW + (r + 2); }
  }
  const int mslot = (ty + 1) * PW + (tx + 1);

  const T dt2 = -0.5 * dt;
  const T dtx = dt * P.dxinv, dty = dt * P.dyinv, dtz = dt * P.dzinv;

  T wm[5], wc[5], wp[5], qn[6], qc[6], tmpq[6];
  T part[6], zr_prev[5], fz_prev[6];
  double vmin = 1e6;

  load_cell<S, T>(P, G, qin, gx, gy, z0 - 2, tmpq, wm);
  load_cell<S, T>(P, G, qin, gx, gy, z0 - 1, tmpq, wc);
#pragma unroll
  for (int m = 0; m < 6; m++) { part[m] = 0; fz_prev[m] = 0; qc[m] = 0; }
#pragma unroll
  for (int m = 0; m < 5; m++) zr_prev[m] = 0;

  for (int k = z0 - 1; k <= z1; k++) {
    // 1. primitive of plane k+1 for this column
    load_cell<S, T>(P, G, qin, gx, gy, k + 1, qn, wp);
    // 2. ring cells of plane k
    if (rslot >= 0) {
      T rq[6], rw[5];
      load_cell<S, T>(P, G, qin, rgx, rgy, k, rq, rw);
#pragma unroll
      for (int m = 0; m < 5; m++) wpl[m][rslot] = rw[m];
    }
#pragma unroll
    for (int m = 0; m < 5; m++) wpl[m][mslot] = wc[m];
    __syncthreads();

    // 3. slopes and predictor of plane k (muscl.f90:93-107)
    T sx[5], sy[5], sz[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sx[m] = slope3<LIMITER>(wpl[m][mslot - 1], wc[m], wpl[m][mslot + 1]);
      sy[m] = slope3<LIMITER>(wpl[m][mslot - PW], wc[m], wpl[m][mslot + PW]);
      sz[m] = slope3<LIMITER>(wm[m], wc[m], wp[m]);
    }
    T ev[5];
    if (P.secondorder) {
      T a1[5], a2[5], a3[5];
      predictor_dir<0>(P, wc, sx, a1);
      predictor_dir<1>(P, wc, sy, a2);
      predictor_dir<2>(P, wc, sz, a3);
#pragma unroll
      for (int m = 0; m < 5; m++) ev[m] = dt2 * (a1[m] * P.dxinv + a2[m] * P.dyinv + a3[m] * P.dzinv);
    } else {
#pragma unroll
      for (int m = 0; m < 5; m++) { ev[m] = 0; sx[m] = 0; sy[m] = 0; sz[m] = 0; }
    }
    // face states (muscl.f90:207-210): right/top = w + 0.5 s + evo, left/bottom = w - 0.5 s + evo
    T xl[5], yl[5], zl[5], zr[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      xs[m][tid] = wc[m] + 0.5 * sx[m] + ev[m];
      ys[m][tid] = wc[m] + 0.5 * sy[m] + ev[m];
      xl[m] = wc[m] - 0.5 * sx[m] + ev[m];
      yl[m] = wc[m] - 0.5 * sy[m] + ev[m];
      zl[m] = wc[m] - 0.5 * sz[m] + ev[m];
      zr[m] = wc[m] + 0.5 * sz[m] + ev[m];
    }
    __syncthreads();

    const bool out_plane = k >= z0 && k < z1;
    // 4. x and y faces of plane k
    if (out_plane) {
      if (tx >= 1 && ty >= 1 && ty <= TY) {
        T wl[5], f[6];
#pragma unroll
        for (int m = 0; m < 5; m++) wl[m] = xs[m][tid - 1];
        const int e = riemann<SOLVER, 1>(P, wl, xl, f);
        if (e && gx <= G.nx && gy < G.ny) report(st, e, (unsigned long long)(gx + (long long)G.nx * (gy + (long long)G.ny * k)));
#pragma unroll
        for (int m = 0; m < 6; m++) fx[m][tid] = f[m];
      }
      if (ty >= 1 && tx >= 1 && tx <= TX) {
        T wl[5], f[6];
#pragma unroll
        for (int m = 0; m < 5; m++) wl[m] = ys[m][tid - BX];
        const int e = riemann<SOLVER, 2>(P, wl, yl, f);
        if (e && gx < G.nx && gy <= G.ny) report(st, e, (unsigned long long)(gx + (long long)G.nx * (gy + (long long)G.ny * k)));
#pragma unroll
        for (int m = 0; m < 6; m++) fy[m][tid] = f[m];
      }
    }
    __syncthreads();

    if (interior) {
      // 5. z face k-1/2, finish plane k-1
      if (k >= z0) {
        T f[6];
        const int e = riemann<SOLVER, 3>(P, zr_prev, zl, f);
        if (e) report(st, e, (unsigned long long)(gx + (long long)G.nx * (gy + (long long)G.ny * k)));
        if (k - 1 >= z0) {
          T o[6];
#pragma unroll
          for (int m = 0; m < 6; m++) o[m] = part[m] + dtz * (fz_prev[m] - f[m]);
          const long long c = (long long)gx + (long long)G.nx * ((long long)gy + (long long)G.ny * (k - 1));
#pragma unroll
          for (int m = 0; m < 6; m++) qout[m * G.n + c] = S(o[m]);
          // primitive + CFL of the new state (next step's primitive/hydrostep)
          T onew[6], wn[5];
#pragma unroll
          for (int m = 0; m < 6; m++) onew[m] = T(S(o[m]));
          const int ep = q2w(P, onew, wn);
          if (ep) report_next(st, ep, (unsigned long long)c);
          vmin = dmin(vmin, cfl_cell(P, wn));
        }
#pragma unroll
        for (int m = 0; m < 6; m++) fz_prev[m] = f[m];
      }
      // 6. x then y differences of plane k
      if (out_plane) {
#pragma unroll
        for (int m = 0; m < 6; m++) part[m] = qc[m] + dtx * (fx[m][tid] - fx[m][tid + 1]);
#pragma unroll
        for (int m = 0; m < 6; m++) part[m] = part[m] + dty * (fy[m][tid] - fy[m][tid + BX]);
      }
#pragma unroll
      for (int m = 0; m < 5; m++) zr_prev[m] = zr[m];
    }
    // 7. rotate the column window
#pragma unroll
    for (int m = 0; m < 5; m++) { wm[m] = wc[m]; wc[m] = wp[m]; }
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = qn[m];
  }

  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((tid & 63) == 0) atomic_min_pos(&st->vmin_bits, vmin);
}

// primitive + CFL min over q without materialising w (first step, after uploads)
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
  g.zchunk = 64;
  if (g.zchunk > g.nz) g.zchunk = g.nz;
  g.tiles_x = (g.nx + TX - 1) / TX;
  return g;
}

template <int SOLVER, int LIMITER>
void launch_fused_t(evh_ctx *h, double dt) {
  const Geo g = make_geo(h);
  const int tiles_y = (g.ny + TY - 1) / TY;
  const dim3 grid(g.tiles_x * tiles_y, (g.nz + g.zchunk - 1) / g.zchunk);
  const double *qin = (const double *)h->qbuf[h->cur];
  double *qout = (double *)h->qbuf[1 - h->cur];
  hipLaunchKernelGGL((k_fused<SOLVER, LIMITER, double>), grid, dim3(NT), 0, h->stream, h->kp, g, qin, qout, dt, h->st);
}

}  // namespace

bool fused_supported(const evh_config &c) {
  return c.nx >= 2 && c.ny >= 2 && c.nz >= 2 && c.precision == EVH_FP64;
}

hipError_t launch_cfl(evh_ctx *h) {
  const Geo g = make_geo(h);
  long long b = (g.n + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_cfl<double>, dim3((int)b), dim3(256), 0, h->stream, h->kp, g, (const double *)h->qbuf[h->cur],
                     h->st);
  return hipGetLastError();
}

hipError_t launch_fused_step(evh_ctx *h, double dt) {
  const int s = h->kp.solver, l = h->kp.limiter;
  if (s == 1 && l == 1) launch_fused_t<1, 1>(h, dt);
  else if (s == 1 && l == 2) launch_fused_t<1, 2>(h, dt);
  else if (s == 1 && l == 3) launch_fused_t<1, 3>(h, dt);
  else if (s == 2 && l == 1) launch_fused_t<2, 1>(h, dt);
  else if (s == 2 && l == 2) launch_fused_t<2, 2>(h, dt);
  else if (s == 2 && l == 3) launch_fused_t<2, 3>(h, dt);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace evh
