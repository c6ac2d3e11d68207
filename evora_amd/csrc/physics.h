// physics.h -- per-cell and per-face arithmetic of evora's hydro step, as
// device functions shared by every kernel of libevora_hip.
//
// Operand order follows the reference expression by expression, so that a
// translation unit built with -ffp-contract=off reproduces the reference's
// IEEE double results (the primitive and CFL functions, which the fast
// translation units call for the first step's limit, turn contraction off
// themselves) (the staged kernels are checked bit for bit against
// the CPU restatement).  Fused kernels include the same code; where they
// trade exactness for speed it is a named alternative below, not an edit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace evh {

// error codes (mirror include/evora_hip.h)
enum : int {
  kOk = 0,
  kErrNegPressure = 1,
  kErrRiemannPressure = 2,
  kErrHll = 3,
  kErrHllc = 4,
  kErrDt = 5,
  kErrSolver = 6,
  kErrLimiter = 7,
};

// Constants of module `global` the step reads (global.f90:44-86).
struct KParams {
  int nx, ny, nz;
  int boundtype, solver, limiter, secondorder, dualenergy;
  int onedimension, twodimension;
  double g, g1, g1inv, gg1, cfl, de_thresh1;
  double dx, dy, dz, dxinv, dyinv, dzinv;
  double mach_fp64;
  double kin_frac64;  // FP32-mixed: kin/E above which p is recovered in FP64
  double g1_dln2;     // (2/3 - g1) ln 2
  int g1_third;       // |2/3 - g1| <= 1e-6: powers of g1 via the cube root
  int conduction;     // thermal conduction (simplekappa) on
  double c_cond, temp0, kappa0;
};

// Device status block: first error (code, linear cell) and min reductions.
struct Status {
  int code;           // first error code seen (0 = none)
  int pad;
  unsigned long long cell;  // linear index of an offending cell (x fastest), or ~0
  unsigned long long vmin_bits;   // CFL min of dx/(|u|+a) (positive double bits)
  unsigned long long next_code;   // error found while computing the next step's primitive
  unsigned long long next_cell;
  unsigned long long tvmin_bits;  // conduction limit: min of rho / kappa (positive double bits)
  unsigned long long pad_;
};

__device__ __forceinline__ void report(Status *st, int code, unsigned long long cell) {
  if (atomicCAS(&st->code, 0, code) == 0) atomicExch(&st->cell, cell);
}

__device__ __forceinline__ void report_next(Status *st, int code, unsigned long long cell) {
  if (atomicCAS(&st->next_code, 0ull, (unsigned long long)code) == 0ull) atomicExch(&st->next_cell, cell);
}

// positive doubles order like their bit patterns
__device__ __forceinline__ void atomic_min_pos(unsigned long long *a, double v) {
  atomicMin(a, (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }

// ---------------------------------------------------------------- comp.f90:45-85
// pressure(q, press); q may be rewritten when dualenergy is on.
template <typename T>
__device__ __forceinline__ int pressure(const KParams &P, T q[6], T &press) {
#pragma clang fp contract(off)
  const T m2 = (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]);
  const T kin = T(0.5) * m2 / q[0];
  const T ratio = kin / q[4];
  T pr;
  if (P.dualenergy) {
    if (ratio > T(P.de_thresh1)) {
      pr = q[5] * pow(q[0], T(P.g1));
      q[4] = kin + pr * T(P.g1inv);
    } else {
      pr = T(P.g1) * (q[4] - kin);
      q[5] = pr / pow(q[0], T(P.g1));
    }
  } else {
    pr = T(P.g1) * (q[4] - kin);
  }
  press = pr;
  return (pr <= T(0)) ? kErrNegPressure : kOk;
}

// comp.f90:26-43  q2w
template <typename T>
__device__ __forceinline__ int q2w(const KParams &P, T q[6], T w[5]) {
#pragma clang fp contract(off)
  const T densinv = T(1) / q[0];
  w[0] = q[0];
  w[1] = q[1] * densinv;
  w[2] = q[2] * densinv;
  w[3] = q[3] * densinv;
  return pressure(P, q, w[4]);
}

// timestep.f90:39-44: min over directions of d/(|u|+a) for one cell
template <typename T>
__device__ __forceinline__ double cfl_cell(const KParams &P, const T w[5]) {
#pragma clang fp contract(off)
  const double a = sqrt(P.g * double(w[4]) / double(w[0]));
  const double vx = P.dx / (fabs(double(w[1])) + a);
  const double vy = P.dy / (fabs(double(w[2])) + a);
  const double vz = P.dz / (fabs(double(w[3])) + a);
  return dmin(dmin(vx, vy), vz);
}

// ---------------------------------------------------------------- muscl.f90:115-169
// one component of comp_slope from the one-sided differences
template <int LIMITER, typename T>
__device__ __forceinline__ T limited_slope(T dwl, T dwr) {
  if (LIMITER == 1) {  // minmod
    const bool lpos = dwl >= T(0), rpos = dwr >= T(0);
    if (lpos != rpos) return T(0);
    return rpos ? dmin(dwl, dwr) : dmax(dwl, dwr);
  } else if (LIMITER == 2) {  // monotonized central
    const T dwc = (dwr + dwl) * T(0.5);
    const bool lpos = dwl >= T(0), rpos = dwr >= T(0);
    if (lpos != rpos) return T(0);
    return rpos ? dmin(dmin(T(2) * dwl, dwc), T(2) * dwr) : dmax(dmax(T(2) * dwl, dwc), T(2) * dwr);
  } else {  // van Leer
    const T pro = dwl * dwr;
    if (pro <= T(0)) return T(0);
    return T(2) * pro / (dwr + dwl);
  }
}

// MC needs (wr - wl)*0.5 exactly as muscl.f90:142; (dwr + dwl) differs in the
// last bit, so the exact variant takes the three states.
template <int LIMITER, typename T>
__device__ __forceinline__ T slope3(T wl, T wc, T wr) {
  const T dwl = wc - wl, dwr = wr - wc;
  if (LIMITER == 2) {
    const T dwc = (wr - wl) * T(0.5);
    const bool lpos = dwl >= T(0), rpos = dwr >= T(0);
    if (lpos != rpos) return T(0);
    return rpos ? dmin(dmin(T(2) * dwl, dwc), T(2) * dwr) : dmax(dmax(T(2) * dwl, dwc), T(2) * dwr);
  }
  return limited_slope<LIMITER>(dwl, dwr);
}

// muscl.f90:100-104: predictor contribution of direction D (0,1,2)
template <int D, typename T>
__device__ __forceinline__ void predictor_dir(const KParams &P, const T w[5], const T s[5], T a[5]) {
#pragma unroll
  for (int i = 0; i < 5; i++) a[i] = w[D + 1] * s[i];
  a[4] = a[4] + T(P.g) * w[4] * s[D + 1];
  a[D + 1] = a[D + 1] + s[4] / w[0];
}

// ---------------------------------------------------------------- riemann.f90
// flux (riemann.f90:157-205) along DIR (1,2,3); pg = w(1)**g1
template <int DIR, typename T>
__device__ __forceinline__ void flux_pg(const KParams &P, const T w[5], T pg, T f[6]) {
  const T u2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  const T ud = w[DIR];
  f[0] = w[0] * ud;
#pragma unroll
  for (int i = 1; i <= 3; i++) f[i] = (i == DIR) ? (w[0] * (ud * ud) + w[4]) : (w[0] * ud * w[i]);
  f[4] = ud * (T(0.5) * w[0] * u2 + T(P.gg1) * w[4]);
  f[5] = ud * w[4] / pg;
}

// state (riemann.f90:207-239)
template <typename T>
__device__ __forceinline__ void state_pg(const KParams &P, const T w[5], T pg, T q[6]) {
  q[0] = w[0];
  q[1] = w[0] * w[1];
  q[2] = w[0] * w[2];
  q[3] = w[0] * w[3];
  const T u2 = (w[1] * w[1]) + (w[2] * w[2]) + (w[3] * w[3]);
  q[4] = T(0.5) * w[0] * u2 + w[4] / T(P.g1);
  q[5] = w[4] / pg;
}

// statex (riemann.f90:241-288); pgg = w(1)**g
template <int DIR, typename T>
__device__ __forceinline__ void statex_pg(const KParams &P, const T w[5], T pgg, T s, T sm, T qx[6]) {
  const T ud = w[DIR];
  qx[0] = w[0] * (s - ud) / (s - sm);
#pragma unroll
  for (int i = 1; i <= 3; i++) qx[i] = ((i == DIR) ? sm : w[i]) * qx[0];
  const T energy = T(0.5) * w[0] * ((w[1] * w[1]) + (w[2] * w[2]) + (w[3] * w[3])) + w[4] / T(P.g1);
  qx[4] = qx[0] * (energy / w[0] + (sm - ud) * (sm + (w[4] / (w[0] * (s - ud)))));
  qx[5] = qx[0] * w[4] / pgg;
}

// hll (riemann.f90:47-92)
template <int DIR, typename T>
__device__ __forceinline__ int hll(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  if (wl[4] <= T(0) || wr[4] <= T(0)) return kErrRiemannPressure;
  const T al = sqrt(T(P.g) * wl[4] / wl[0]);
  const T ar = sqrt(T(P.g) * wr[4] / wr[0]);
  const T sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const T sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  if (T(0) < sl) {
    flux_pg<DIR>(P, wl, T(pow(wl[0], T(P.g1))), f);
  } else if (T(0) > sr) {
    flux_pg<DIR>(P, wr, T(pow(wr[0], T(P.g1))), f);
  } else if (sl <= T(0) && T(0) <= sr) {
    T fl[6], fr[6], ql[6], qr[6];
    const T pgl = pow(wl[0], T(P.g1)), pgr = pow(wr[0], T(P.g1));
    flux_pg<DIR>(P, wl, pgl, fl);
    flux_pg<DIR>(P, wr, pgr, fr);
    state_pg(P, wl, pgl, ql);
    state_pg(P, wr, pgr, qr);
#pragma unroll
    for (int i = 0; i < 6; i++) f[i] = (sr * fl[i] - sl * fr[i] + sl * sr * (qr[i] - ql[i])) / (sr - sl);
  } else {
    return kErrHll;
  }
  return kOk;
}

// hllc (riemann.f90:96-153): only the selected side's flux and star state
template <int DIR, typename T>
__device__ __forceinline__ int hllc(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  if (wl[4] <= T(0) || wr[4] <= T(0)) return kErrRiemannPressure;
  const T al = sqrt(T(P.g) * wl[4] / wl[0]);
  const T ar = sqrt(T(P.g) * wr[4] / wr[0]);
  const T sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const T sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  const T facl = wl[0] * (sl - wl[DIR]);
  const T facr = wr[0] * (sr - wr[DIR]);
  const T sm = (wr[4] - wl[4] + facl * wl[DIR] - facr * wr[DIR]) / (facl - facr);
  // branch: 0 left, 1 right, 2 left star, 3 right star
  int br;
  if (T(0) < sl) br = 0;
  else if (T(0) > sr) br = 1;
  else if (sl < T(0) && T(0) < sm) br = 2;
  else if (sm <= T(0) && T(0) < sr) br = 3;
  else return kErrHllc;
  const bool right = (br == 1 || br == 3);
  T w[5];
#pragma unroll
  for (int i = 0; i < 5; i++) w[i] = right ? wr[i] : wl[i];
  const T pg = pow(w[0], T(P.g1));
  flux_pg<DIR>(P, w, pg, f);
  if (br >= 2) {
    const T s = right ? sr : sl;
    T q[6], qx[6];
    state_pg(P, w, pg, q);
    statex_pg<DIR>(P, w, T(pow(w[0], T(P.g))), s, sm, qx);
#pragma unroll
    for (int i = 0; i < 6; i++) f[i] = f[i] + s * (qx[i] - q[i]);
  }
  return kOk;
}

template <int SOLVER, int DIR, typename T>
__device__ __forceinline__ int riemann(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
  if (SOLVER == 1) return hll<DIR>(P, wl, wr, f);
  return hllc<DIR>(P, wl, wr, f);
}

// ---------------------------------------------------------------- boundary.f90
// Ghost index map of `bound` for one axis: index g in [-2, n+1] -> source
// cell in [0, n).  Zero-gradient ghosts mirror (w(lx-2) = w(lx+1),
// boundary.f90:36-37); periodic wraps.  Composing the per-axis maps
// reproduces the x -> y -> z fill over full extents (edges and corners).
__device__ __forceinline__ int ghost_src(int g, int n, bool periodic) {
  if (g < 0) return periodic ? g + n : -1 - g;
  if (g >= n) return periodic ? g - n : 2 * n - 1 - g;
  return g;
}

}  // namespace evh
