// physics_fast.h -- the per-cell and per-face arithmetic of physics.h in the
// form the fused kernels execute: every division by a value that is used
// more than once becomes one refined reciprocal, rho**gamma is taken as
// rho * rho**(gamma-1) so that each face evaluates one power, and the
// entropy flux reuses p / rho**(gamma-1) from the state.  Same formulas and
// branch structure as the reference (comp.f90:45-85, muscl.f90:100-104,
// riemann.f90:47-288); results agree with the exact forms to a few ulp.
#pragma once

#include "fastmath.h"
#include "physics.h"

namespace evh {
namespace fast {

// rho**(-g1) and rho**g1 of a positive density: the cube-root form of
// fastmath.h when g1 is close to 2/3 and rho is inside 2^(+-200), the
// general exp(y ln x) otherwise; single precision uses the hardware
// exp2/log2.
// the general power, out of line: the cube-root form covers the hot cases,
// and inlined here the series constants would occupy registers for the
// whole of a kernel's loop
__device__ __attribute__((noinline)) double pow_general(double x, double y) { return fm::pow_pos(x, y); }

__device__ __forceinline__ double rpow_g1(const KParams &P, double x) {
  const int e = __builtin_amdgcn_frexp_exp(x);
  if (P.g1_third && e > -200 && e < 200) {
    double y, c;
    fm::third_root(x, e, P.g1_dln2, y, c);
    const double z = y * y;
    return __builtin_fma(z, fm::expm1_tiny(c), z);
  }
  return pow_general(x, -P.g1);
}
template <bool HALLEY = true>
__device__ __forceinline__ double pow_g1(const KParams &P, double x) {
  const int e = __builtin_amdgcn_frexp_exp(x);
  if (P.g1_third && e > -200 && e < 200) {
    double y, c;
    fm::third_root<HALLEY>(x, e, P.g1_dln2, y, c);
    const double v = x * y;
    return __builtin_fma(v, fm::expm1_tiny(-c), v);
  }
  return pow_general(x, P.g1);
}
__device__ __forceinline__ float rpow_g1(const KParams &P, float x) { return fm::pow_pos(x, -float(P.g1)); }
__device__ __forceinline__ float pow_g1(const KParams &P, float x) { return fm::pow_pos(x, float(P.g1)); }

// comp.f90:26-85 (q2w + pressure); q may be rewritten when dualenergy is on.
// Returns 1/rho in rinv.
//
// FP64: the reference's roundings.  The pressure p = g1 (E - kin) feeds the
// normal-momentum flux, whose difference across a cell is the whole
// momentum update wherever the flow along that axis is slow; one ulp of p
// there is a relative error of ulp(p) / |rho u_d|.  So densinv = 1/rho, the
// velocities, m2 (no fused multiply-add), kin = 0.5 m2 / rho and
// ratio = kin/E are formed as the reference forms them, the two divisions
// by rho sharing one reciprocal (div_cr); only the powers of the dual-energy
// branches take the fast form.
__device__ __forceinline__ int q2w(const KParams &P, double q[6], double w[5], double &rinv) {
#pragma clang fp contract(off)
  const double y = fm::rcp(q[0]);
  const double densinv = __builtin_fma(__builtin_fma(-q[0], y, 1.0), y, y);  // 1.0 / q(1)
  rinv = densinv;
  w[0] = q[0];
  w[1] = q[1] * densinv;
  w[2] = q[2] * densinv;
  w[3] = q[3] * densinv;
  const double m2 = (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]);
  const double kin = fm::div_cr(0.5 * m2, q[0], y);
  double pr;
  if (P.dualenergy) {
    if (fm::div_cr(kin, q[4], fm::rcp(q[4])) > P.de_thresh1) {  // ratio = kin/E > de_thresh1
      pr = q[5] * pow_g1(P, q[0]);
      q[4] = kin + pr * P.g1inv;
    } else {
      pr = P.g1 * (q[4] - kin);
      q[5] = pr * rpow_g1(P, q[0]);
    }
  } else {
    pr = P.g1 * (q[4] - kin);
  }
  w[4] = pr;
  return (pr <= 0.0) ? kErrNegPressure : kOk;
}

// FP32: refined hardware reciprocal, fused arithmetic
template <typename T>
__device__ __forceinline__ int q2w(const KParams &P, T q[6], T w[5], T &rinv) {
  rinv = fm::rcp(q[0]);
  w[0] = q[0];
  w[1] = q[1] * rinv;
  w[2] = q[2] * rinv;
  w[3] = q[3] * rinv;
  const T m2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const T kin = T(0.5) * m2 * rinv;
  T pr;
  if (P.dualenergy) {
    if (kin * fm::rcp(q[4]) > T(P.de_thresh1)) {  // ratio = kin/E > de_thresh1
      pr = q[5] * pow_g1(P, q[0]);
      q[4] = kin + pr * T(P.g1inv);
    } else {
      pr = T(P.g1) * (q[4] - kin);
      q[5] = pr * rpow_g1(P, q[0]);
    }
  } else {
    pr = T(P.g1) * (q[4] - kin);
  }
  w[4] = pr;
  return (pr <= T(0)) ? kErrNegPressure : kOk;
}

// timestep.f90:39-44 for one cell, from w and 1/rho:
// min_d dx_d / (|u_d| + a) taken as 1 / max_d (|u_d| + a) / dx_d, one
// reciprocal instead of three (equal to a few ulp).
template <typename T>
__device__ __forceinline__ double cfl_speed(const KParams &P, const T w[5], T rinv) {
  const double a = fm::sqrt_pos(P.g * double(w[4]) * double(rinv));
  const double sx = (fabs(double(w[1])) + a) * P.dxinv;
  const double sy = (fabs(double(w[2])) + a) * P.dyinv;
  const double sz = (fabs(double(w[3])) + a) * P.dzinv;
  return dmax(dmax(sx, sy), sz);
}
template <typename T>
__device__ __forceinline__ double cfl_cell(const KParams &P, const T w[5], T rinv) {
  return fm::rcp(cfl_speed(P, w, rinv));
}

// muscl.f90:100-104.  FP64 in the reference's roundings (the predicted
// pressure enters the face states like the primitive one): s(5)/rho is the
// IEEE quotient from the correctly rounded rinv; FP32 as s(5)*rinv.
template <int D>
__device__ __forceinline__ void predictor_dir(const KParams &P, const double w[5], double rinv, const double s[5],
                                              double a[5]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 5; i++) a[i] = w[D + 1] * s[i];
  a[4] = a[4] + P.g * w[4] * s[D + 1];
  a[D + 1] = a[D + 1] + fm::div_cr(s[4], w[0], rinv);
}
template <int D>
__device__ __forceinline__ void predictor_dir(const KParams &P, const float w[5], float rinv, const float s[5],
                                              float a[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) a[i] = w[D + 1] * s[i];
  a[4] = fm::fmad(float(P.g) * w[4], s[D + 1], a[4]);
  a[D + 1] = fm::fmad(s[4], rinv, a[D + 1]);
}

// evo (muscl.f90:106-107) from the three directions' predictor terms:
// FP64 as the reference, dt2 (a1 dxinv + a2 dyinv + a3 dzinv) unfused; FP32
// with dt2 folded into the three factors
__device__ __forceinline__ void evo5(const KParams &P, double dt2, const double a1[5], const double a2[5],
                                     const double a3[5], double ev[5]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int m = 0; m < 5; m++) ev[m] = dt2 * (a1[m] * P.dxinv + a2[m] * P.dyinv + a3[m] * P.dzinv);
}
// the 2-D form (muscl.f90:79-85): dt2 (a1 dxinv + a2 dyinv)
__device__ __forceinline__ void evo5_2d(const KParams &P, double dt2, const double a1[5], const double a2[5],
                                        double ev[5]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int m = 0; m < 5; m++) ev[m] = dt2 * (a1[m] * P.dxinv + a2[m] * P.dyinv);
}
__device__ __forceinline__ void evo5_2d(const KParams &P, float dt2, const float a1[5], const float a2[5],
                                        float ev[5]) {
  const float ex = dt2 * float(P.dxinv), ey = dt2 * float(P.dyinv);
#pragma unroll
  for (int m = 0; m < 5; m++) ev[m] = fm::fmad(a1[m], ex, a2[m] * ey);
}
__device__ __forceinline__ void evo5(const KParams &P, float dt2, const float a1[5], const float a2[5],
                                     const float a3[5], float ev[5]) {
  const float ex = dt2 * float(P.dxinv), ey = dt2 * float(P.dyinv), ez = dt2 * float(P.dzinv);
#pragma unroll
  for (int m = 0; m < 5; m++) ev[m] = fm::fmad(a1[m], ex, fm::fmad(a2[m], ey, a3[m] * ez));
}

// muscl.f90:115-169, one component.  minmod and monotonized central in
// the branch-free form max(min(..), min(0, max(..))): it is the smallest
// candidate when all are positive, the largest when all are negative, and
// 0 when the one-sided differences disagree in sign, as in the reference.
template <int LIMITER, typename T>
__device__ __forceinline__ T slope3(T wl, T wc, T wr) {
  const T dwl = wc - wl, dwr = wr - wc;
  if (LIMITER == 1) {
    return fmax(fmin(dwl, dwr), fmin(T(0), fmax(dwl, dwr)));
  } else if (LIMITER == 2) {
    const T dwc = (wr - wl) * T(0.5);
    const T a = T(2) * dwl, b = T(2) * dwr;
    return fmax(fmin(fmin(a, dwc), b), fmin(T(0), fmax(fmax(a, dwc), b)));
  } else {
    const T pro = dwl * dwr;
    if (pro <= T(0)) return T(0);
    return T(2) * pro * fm::rcp(dwr + dwl);
  }
}

// |u|^2 = (u_1^2 + u_2^2) + u_3^2 with the first two squares rounded on
// their own: the same bits when u_1 and u_2 are exchanged (the ring wave
// solves a y face as an x face), whatever contraction the compiler chooses
// around it.  (The FP32 build of the march is compiled without contraction.)
__device__ __forceinline__ double speed2(const double w[5]) {
  return __builtin_fma(w[3], w[3], __dmul_rn(w[1], w[1]) + __dmul_rn(w[2], w[2]));
}
__device__ __forceinline__ float speed2(const float w[5]) { return __builtin_fmaf(w[3], w[3], w[1] * w[1] + w[2] * w[2]); }

// flux (riemann.f90:157-205) of one state, with sq = p / rho**g1
template <int DIR, typename T>
__device__ __forceinline__ void flux(const KParams &P, const T w[5], T u2, T sq, T f[6]) {
  const T ud = w[DIR];
  const T mu = w[0] * ud;
  f[0] = mu;
#pragma unroll
  for (int i = 1; i <= 3; i++) f[i] = (i == DIR) ? fm::fmad(mu, ud, w[4]) : (mu * w[i]);
  f[4] = ud * fm::fmad(T(0.5) * w[0], u2, T(P.gg1) * w[4]);
  f[5] = ud * sq;
}

// state (riemann.f90:207-239)
template <typename T>
__device__ __forceinline__ void state(const KParams &P, const T w[5], T u2, T sq, T q[6]) {
  q[0] = w[0];
  q[1] = w[0] * w[1];
  q[2] = w[0] * w[2];
  q[3] = w[0] * w[3];
  q[4] = fm::fmad(T(0.5) * w[0], u2, w[4] * T(P.g1inv));
  q[5] = sq;
}

// hll (riemann.f90:47-92), branch-free: both sides' flux and state are
// formed and the region (0 < sl: left, 0 > sr: right, else the HLL
// average) is selected, so that the solver is one basic block and
// independent solves can be interleaved by the compiler.  The error codes
// are the reference's: non-positive pressure first, then no region.
template <int DIR, typename T>
__device__ __forceinline__ int hll(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  int code = (wl[4] <= T(0) || wr[4] <= T(0)) ? kErrRiemannPressure : kOk;
  const T ril = fm::rcp(wl[0]), rir = fm::rcp(wr[0]);
  const T al = fm::sqrt_pos(T(P.g) * wl[4] * ril);
  const T ar = fm::sqrt_pos(T(P.g) * wr[4] * rir);
  const T sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const T sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  const T u2l = speed2(wl);
  const T u2r = speed2(wr);
  const bool left = T(0) < sl;
  const bool right = T(0) > sr;
  const bool mid = sl <= T(0) && T(0) <= sr;
  if (!left && !right && !mid && !code) code = kErrHll;
  T fl[6], fr[6], ql[6], qr[6];
  const T sql = wl[4] * rpow_g1(P, wl[0]);
  const T sqr = wr[4] * rpow_g1(P, wr[0]);
  fast::flux<DIR>(P, wl, u2l, sql, fl);
  fast::flux<DIR>(P, wr, u2r, sqr, fr);
  fast::state(P, wl, u2l, sql, ql);
  fast::state(P, wr, u2r, sqr, qr);
  const T inv = fm::rcp(sr - sl);
  const T slsr = sl * sr;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    T fm_;
    if constexpr (sizeof(T) == 4) fm_ = fm::fmad(slsr, qr[i] - ql[i], fm::fmad(sr, fl[i], -(sl * fr[i]))) * inv;
    else fm_ = (sr * fl[i] - sl * fr[i] + slsr * (qr[i] - ql[i])) * inv;
    f[i] = left ? fl[i] : (right ? fr[i] : fm_);
  }
  return code;
}

// sound speed sqrt(g p / rho) = g p / sqrt(g p rho)
template <typename T>
__device__ __forceinline__ T sound(const KParams &P, T rho, T p) {
  const T gp = T(P.g) * p;
  return gp * fm::rsqrt_pos(gp * rho);
}

// hllc (riemann.f90:96-153): the flux and star state of the selected side
// only.  statex (riemann.f90:241-288) is used in the rearranged form
//   r = (s-u)/(s-sm), qx1 = rho r, qx5 = r E + (sm-u)(qx1 sm + p/(s-sm)),
//   qx6 = r p/rho**g1,
// which equals the reference expression and needs no 1/rho or 1/(s-u).
template <int DIR, typename T>
__device__ __forceinline__ int hllc(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  // branch-free, like hll: the region is an index, the star correction is
  // always formed and kept only in the star regions
  int code = (wl[4] <= T(0) || wr[4] <= T(0)) ? kErrRiemannPressure : kOk;
  const T al = sound(P, wl[0], wl[4]);
  const T ar = sound(P, wr[0], wr[4]);
  const T sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const T sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  const T facl = wl[0] * (sl - wl[DIR]);
  const T facr = wr[0] * (sr - wr[DIR]);
  T sm;
  if constexpr (sizeof(T) == 4)
    sm = fm::fmad(-facr, wr[DIR], fm::fmad(facl, wl[DIR], wr[4] - wl[4])) * fm::rcp(facl - facr);
  else
    sm = (wr[4] - wl[4] + facl * wl[DIR] - facr * wr[DIR]) * fm::rcp(facl - facr);
  // FP64: the reference's regions, with its error where none holds (a wave
  // speed exactly 0: sl = 0 with sm > 0, or sr = 0 with sm <= 0).  FP32
  // rounds onto such a speed far more often (a transonic face every few
  // million in a Mach 17 cloud); there the star and outer fluxes coincide,
  // so the FP32 form takes the star region and stops only on a NaN speed.
  int br;
  if constexpr (sizeof(T) == 8)
    br = (T(0) < sl) ? 0 : (T(0) > sr) ? 1 : (sl < T(0) && T(0) < sm) ? 2 : (sm <= T(0) && T(0) < sr) ? 3 : -1;
  else
    br = (T(0) < sl) ? 0 : (T(0) > sr) ? 1 : (T(0) < sm) ? 2 : (sm <= T(0)) ? 3 : -1;
  if (br < 0 && !code) code = kErrHllc;
  const bool right = br == 1 || br == 3;
  T w[5];
#pragma unroll
  for (int i = 0; i < 5; i++) w[i] = right ? wr[i] : wl[i];
  const T u2 = speed2(w);
  const T sq = w[4] * rpow_g1(P, w[0]);  // p / rho**g1
  const T ud = w[DIR];
  const T mu = w[0] * ud;
  const T energy = fm::fmad(T(0.5) * w[0], u2, w[4] * T(P.g1inv));
  f[0] = mu;
#pragma unroll
  for (int i = 1; i <= 3; i++) f[i] = (i == DIR) ? fm::fmad(mu, ud, w[4]) : (mu * w[i]);
  f[4] = ud * (energy + w[4]);  // ud (0.5 rho u2 + gg1 p) with gg1 = 1 + 1/g1
  f[5] = ud * sq;
  const T s = right ? sr : sl;
  const bool star = br >= 2;
  // outside the star regions the correction is multiplied by 0; its
  // divisor is kept finite there
  const T isd = fm::rcp(star ? s - sm : T(1));
  const T r = (s - ud) * isd;
  const T qx0 = w[0] * r;
  T dq[6];
  dq[0] = qx0 - w[0];
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int i = 1; i <= 3; i++) dq[i] = fm::fmad((i == DIR) ? sm : w[i], qx0, -(w[0] * w[i]));
    dq[4] = fm::fmad(r, energy, fm::fmad(sm - ud, fm::fmad(qx0, sm, w[4] * isd), -energy));
    dq[5] = fm::fmad(r, sq, -sq);
  } else {
#pragma unroll
    for (int i = 1; i <= 3; i++) dq[i] = ((i == DIR) ? sm : w[i]) * qx0 - w[0] * w[i];
    dq[4] = r * energy + (sm - ud) * (qx0 * sm + w[4] * isd) - energy;
    dq[5] = r * sq - sq;
  }
  const T ss = star ? s : T(0);
#pragma unroll
  for (int i = 0; i < 6; i++) f[i] = fm::fmad(ss, dq[i], f[i]);
  return code;
}

template <int SOLVER, int DIR, typename T>
__device__ __forceinline__ int riemann(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  if (SOLVER == 1) return fast::hll<DIR>(P, wl, wr, f);
  return fast::hllc<DIR>(P, wl, wr, f);
}

}  // namespace fast
}  // namespace evh
