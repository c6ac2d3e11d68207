// march2d.h -- the hydro step of a 2-D grid (nz = 1) in one HBM pass: the
// 2-D branch of compevo (muscl.f90:70-88) and the x and y sweeps
// (muscl.f90:171-246), shared arithmetic with the 3-D march (Phys<EXACT>).
//
// A block of NL = 256 lanes owns a strip of NC = 252 columns and marches a
// chunk of rows along y, which plays the part z plays in the 3-D march.
// Lane l holds column x0 + l - 2, so lanes 0-1 and 254-255 hold the two
// ghost columns on either side; the ghost rules of `bound` (boundary.f90)
// are folded into the load index.  Per row k:
//   S1  primitive of row k+1 (its raw q was loaded one row ahead); w of row
//       k to LDS;
//   S2  limited slopes in x (LDS neighbours) and y (rows k-1, k, k+1 in
//       registers), the 2-D predictor, the two states of every x face to
//       LDS; the y face k-1/2 is solved from the top state of row k-1
//       (carried) and the bottom state of row k, and row k-1 is finished as
//       (q + dtx*(Fx)) + dty*(Fy), the reference's x, y order, and stored;
//   S3  the x faces of row k, one per lane;
//   S4  q(k) + dtx*(Fx) of row k, carried to row k+1; the primitive and CFL
//       limit of the stored row k-1 (and its temperature, with conduction).
// Every lane does useful work in every stage except the four ghost lanes in
// the solves and updates, and a cell takes two Riemann solves.
#pragma once

#include "fused_core.h"

namespace evh {
namespace core2d {

constexpr int NL = 256;     // lanes per block
constexpr int NC = NL - 4;  // interior columns per strip

struct Geo2 {
  int nx, ny;
  long long n;
  int ychunk, strips, nchunks, ntiles;
};

// FP64 carries the top face state and bottom flux of row k-1 in lane-private
// LDS rather than in registers: 168 VGPRs and 3 waves per SIMD (52 KB of LDS,
// 3 blocks per CU) instead of 184 and 2.  FP32 keeps them in registers.
template <typename T>
constexpr bool kLdsCarry = sizeof(T) == 8;

template <typename T>
struct Smem2 {
  T W[5][NL];    // w of row k, every lane's column
  T XB[10][NL];  // x face (l, l+1): [0,5) left state, [5,10) right state; then the flux
  T Y[kLdsCarry<T> ? 5 : 1][NL];  // lane-private: the top face state of row k-1
  T F[kLdsCarry<T> ? 6 : 1][NL];  // lane-private: the flux of row k-1's bottom face
};

__device__ __forceinline__ long long col_off(const KParams &P, const Geo2 &G, int gx) {
  const int c = gx < -2 ? -2 : (gx > G.nx + 1 ? G.nx + 1 : gx);
  return ghost_src(c, G.nx, P.boundtype == 2);
}
__device__ __forceinline__ long long row_off(const KParams &P, const Geo2 &G, int gy) {
  const int c = gy < -2 ? -2 : (gy > G.ny + 1 ? G.ny + 1 : gy);
  return (long long)G.nx * ghost_src(c, G.ny, P.boundtype == 2);
}

template <typename S>
__device__ __forceinline__ void fetch2(const Geo2 &G, const S *__restrict__ q, long long c, S raw[6]) {
#pragma unroll
  for (int m = 0; m < 6; m++) raw[m] = q[m * G.n + c];
}

// primitive of a fetched cell of column gx (the reflecting upper x wall of
// boundtype 3, boundary.f90:57-64, negates the velocities of its ghosts)
template <bool EXACT, typename S, typename T>
__device__ __forceinline__ void prim2(const KParams &P, const Geo2 &G, const S raw[6], int gx, T qc[6], T wc[5],
                                      T &rinv) {
#pragma unroll
  for (int m = 0; m < 6; m++) qc[m] = T(raw[m]);
  core::Phys<EXACT>::q2w(P, qc, wc, rinv);
  if (P.boundtype == 3 && gx >= G.nx) {
    wc[1] = -wc[1];
    wc[2] = -wc[2];
    wc[3] = -wc[3];
  }
}

// One strip-chunk `lin` (strips fastest).  As core::march: vmin_dst receives
// the CFL minimum of the new state, errors are reported once after the march,
// and with tout the temperature of each new cell is stored.
template <typename T, typename S, int SOLVER, int LIMITER, bool EXACT>
__device__ __forceinline__ void march2d(const KParams P, const Geo2 G, const S *__restrict__ qin,
                                        S *__restrict__ qout, double dt, Status *st, unsigned long long *vmin_dst,
                                        int lin, Smem2<T> &sm, S *__restrict__ tout) {
  using Ph = core::Phys<EXACT>;
  T(&W)[5][NL] = sm.W;
  T(&XB)[10][NL] = sm.XB;
  constexpr bool LC = kLdsCarry<T>;
  const int l = threadIdx.x;
  const int strip = lin % G.strips, chunk = lin / G.strips;
  const int x0 = strip * NC;
  const int y0 = chunk * G.ychunk, y1 = min(y0 + G.ychunk, G.ny);
  const int gx = x0 + l - 2;
  const bool inner = l >= 2 && l < NL - 2;      // a column of the strip
  const bool inside = inner && gx < G.nx;       // ... inside the grid
  const bool slopes = l >= 1 && l < NL - 1;     // has both x neighbours
  const bool face = l >= 1 && l < NL - 2;       // solves the face (l, l+1)
  const bool face_counts = gx >= -1 && gx < G.nx;  // that face is one the reference solves

  const T dt2 = T(-0.5 * dt);
  const T dtx = T(dt * P.dxinv), dty = T(dt * P.dyinv);
  const T half = T(0.5);

  // rows k-1, k, k+1 of the column; qc, qn: q of rows k and k+1 (the base of
  // the update, kept in registers so that q is read from HBM once)
  T wm[5], wc[5], wp[5], qc[6], qn[6], tq[6];
  T rm, rc, rp;
  // carried from row k-1: its partial update q + dtx*(Fx), the flux of its
  // bottom face, its top face state (the last two in LDS with LC)
  T part[6], fyb[LC ? 1 : 6], ytop[LC ? 1 : 5];
  constexpr bool kSpeedKey = !EXACT && sizeof(T) == 8;
  double vmin = kSpeedKey ? -1e-6 : 1e6;
  int err = 0, err_k = 0, nerr = 0, nerr_k = 0;
  T o[6];
  bool pend = false;
  long long oc = 0;

  S pre[6];
  const long long cx = col_off(P, G, gx);
  {
    S raw[6];
    fetch2(G, qin, cx + row_off(P, G, y0 - 2), raw);
    prim2<EXACT>(P, G, raw, gx, tq, wm, rm);
    fetch2(G, qin, cx + row_off(P, G, y0 - 1), raw);
    prim2<EXACT>(P, G, raw, gx, tq, wc, rc);
  }
  fetch2(G, qin, cx + row_off(P, G, y0), pre);
#pragma unroll
  for (int m = 0; m < 6; m++) { qc[m] = T(0); qn[m] = T(0); part[m] = T(0); }
#pragma unroll
  for (int m = 0; m < 6; m++) {
    if constexpr (LC) sm.F[m][l] = T(0);
    else fyb[m] = T(0);
  }
#pragma unroll
  for (int m = 0; m < 5; m++) {
    if constexpr (LC) sm.Y[m][l] = T(0);
    else ytop[m] = T(0);
  }

  for (int k = y0 - 1; k <= y1; k++) {
    const bool out_row = k >= y0 && k < y1;
    // ---- S1: primitive of row k+1; w of row k to LDS
    prim2<EXACT>(P, G, pre, gx, qn, wp, rp);
    fetch2(G, qin, cx + row_off(P, G, k + 2), pre);
#pragma unroll
    for (int m = 0; m < 5; m++) W[m][l] = wc[m];
    __syncthreads();

    // ---- S2: slopes, predictor, face states; the y face k-1/2
    if (slopes) {
      T sx[5], sy[5], ev[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        sx[m] = Ph::template slope3<LIMITER>(W[m][l - 1], wc[m], W[m][l + 1]);
        sy[m] = Ph::template slope3<LIMITER>(wm[m], wc[m], wp[m]);
      }
      if (P.secondorder) {
        T a1[5], a2[5];
        Ph::template predictor_dir<0>(P, wc, rc, sx, a1);
        Ph::template predictor_dir<1>(P, wc, rc, sy, a2);
        fast::evo5_2d(P, dt2, a1, a2, ev);  // muscl.f90:79-85
      } else {
#pragma unroll
        for (int m = 0; m < 5; m++) { ev[m] = T(0); sx[m] = T(0); sy[m] = T(0); }
      }
      // face states (muscl.f90:207-210): (w + 0.5 s) + evo on the high side,
      // (w - 0.5 s) + evo on the low side
      if (face) {
#pragma unroll
        for (int m = 0; m < 5; m++) XB[m][l] = fm::fmad(half, sx[m], wc[m]) + ev[m];
      }
      if (l >= 2) {
#pragma unroll
        for (int m = 0; m < 5; m++) XB[5 + m][l - 1] = fm::fmad(-half, sx[m], wc[m]) + ev[m];
      }
      if (inner) {
        T yb[5], yt[5];
#pragma unroll
        for (int m = 0; m < 5; m++) {
          yb[m] = fm::fmad(-half, sy[m], wc[m]) + ev[m];
          yt[m] = fm::fmad(half, sy[m], wc[m]) + ev[m];
        }
        if (k >= y0) {
          T fy[6], yt0[5];
#pragma unroll
          for (int m = 0; m < 5; m++) {
            if constexpr (LC) yt0[m] = sm.Y[m][l];
            else yt0[m] = ytop[m];
          }
          const int e = Ph::template riemann<SOLVER, 2>(P, yt0, yb, fy);
          if (e && inside && !err) { err = e; err_k = k; }
          if (k - 1 >= y0 && inside) {
#pragma unroll
            for (int m = 0; m < 6; m++) {
              if constexpr (LC) o[m] = fm::fmad(dty, sm.F[m][l] - fy[m], part[m]);
              else o[m] = fm::fmad(dty, fyb[m] - fy[m], part[m]);
            }
            oc = (long long)gx + (long long)G.nx * (k - 1);
#pragma unroll
            for (int m = 0; m < 6; m++) qout[m * G.n + oc] = S(o[m]);
            pend = true;
          }
#pragma unroll
          for (int m = 0; m < 6; m++) {
            if constexpr (LC) sm.F[m][l] = fy[m];
            else fyb[m] = fy[m];
          }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
          if constexpr (LC) sm.Y[m][l] = yt[m];
          else ytop[m] = yt[m];
        }
      }
    }
    __syncthreads();

    // ---- S3: the x face (l, l+1) of row k
    if (out_row && face) {
      T xl[5], xr[5], fx[6];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        xl[m] = XB[m][l];
        xr[m] = XB[5 + m][l];
      }
      const int e = Ph::template riemann<SOLVER, 1>(P, xl, xr, fx);
      if (e && face_counts && !err) { err = e; err_k = k; }
#pragma unroll
      for (int m = 0; m < 6; m++) XB[m][l] = fx[m];
    }
    __syncthreads();

    // ---- S4: q(k) + dtx*(Fx) of row k; the end of row k-1
    if (out_row && inner) {
#pragma unroll
      for (int m = 0; m < 6; m++) part[m] = fm::fmad(dtx, XB[m][l - 1] - XB[m][l], qc[m]);
    }
    if (pend) {
      T on[6], wn[5], rn;
#pragma unroll
      for (int m = 0; m < 6; m++) on[m] = T(S(o[m]));
      const int ep = Ph::q2w(P, on, wn, rn);
      if (ep && !nerr) { nerr = ep; nerr_k = k - 1; }
      if constexpr (kSpeedKey) vmin = dmin(vmin, -fast::cfl_speed(P, wn, rn));
      else vmin = dmin(vmin, Ph::cfl_cell(P, wn, rn));
      if (tout) tout[oc] = S(Ph::temp(P, wn, rn));
      pend = false;
    }
    // rotate the column window
#pragma unroll
    for (int m = 0; m < 5; m++) { wm[m] = wc[m]; wc[m] = wp[m]; }
    rm = rc;
    rc = rp;
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = qn[m];
  }

  if (err || nerr) {  // the lane's own column (clamped into the grid) locates the error
    const long long ex = min(max(gx, 0), G.nx - 1);
    if (err) report(st, err, (unsigned long long)(ex + G.nx * (long long)err_k));
    if (nerr) report_next(st, nerr, (unsigned long long)(ex + G.nx * (long long)nerr_k));
  }
  if constexpr (kSpeedKey) vmin = fm::rcp(-vmin);
  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((l & 63) == 0) atomic_min_pos(vmin_dst, vmin);
}

}  // namespace core2d
}  // namespace evh
