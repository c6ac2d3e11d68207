// fused_core.h -- the tile march of the fused hydro step, shared by the
// single-grid kernel (fused.hip) and the batched persistent kernel
// (batch.hip).
//
// A block owns a 16 x TY column tile of one box and marches up a chunk of z
// planes, reading q^n once and writing q^{n+1} once; w, evo and s never
// exist in HBM.  Waves 0..TY/4-1 own the 16*TY interior columns one to one;
// the last wave (the "ring wave") owns the 2*(16+TY) columns of the first
// halo ring and loads the second ring and the four corners, so every
// interior lane does useful work in every stage.  Per plane k:
//   S1  primitive (comp.f90:26-85) of plane k+1 for every owned column, with
//       the ghost cells of `bound` (boundary.f90) folded into the load index
//       (a column's x-y offset is formed once per march); plane k goes to
//       LDS with both halo rings.  The ring wave's second-ring cells of
//       plane k were formed in its idle time of S2 of plane k-1.
//   S2  limited slopes in x, y, z and the predictor evo (muscl.f90:45-169)
//       of plane k; the x/y face states go to LDS per face; the interior
//       lanes solve their z face k-1/2 (riemann.f90) and finish plane k-1
//       as (q + dtx*(Fx) + dty*(Fy)) + dtz*(Fz), the reference's x, y, z
//       order (muscl.f90:191-272, integration :287), and store it.
//   S3  x and y faces of plane k: interior lane (i,j) solves its left and
//       bottom faces, the ring wave the TY right-edge and 16 top-edge faces
//       in one solve.  Each flux overwrites the two states it came from.
//   S4  q(k) + dtx*(Fx) + dty*(Fy) of plane k, kept for S2 of plane k+1;
//       then the primitive and the CFL limit of the stored plane k-1, which
//       feed a min reduction for the next hydrostep (timestep.f90:26-57).
//       They sit where the interior waves have the least to do.
// Template parameters: T compute type, S storage type of q (double/double
// for the FP64 form, float/float for the FP32-mixed form), SOLVER, LIMITER,
// TY the tile height (4, 8 or 12).  2-D grids march along y instead
// (march2d.h).
#pragma once

#include "evh_internal.h"
#include "physics_fast.h"

namespace evh {
namespace core {

constexpr int TX = 16;

template <int TY>
struct TileShape {
  static constexpr int NI = TX * TY;          // interior lanes
  static constexpr int NT = NI + 64;          // + the ring wave
  static constexpr int NR1 = 2 * TX + 2 * TY; // first-ring columns
  static constexpr int EW = TX + 4;           // extended plane pitch (x from -2 to TX+1)
  static constexpr int EA = EW * (TY + 4);    // extended plane area
  static constexpr int NFX = (TX + 1) * TY;   // x faces of a plane
  static constexpr int NFY = TX * (TY + 1);   // y faces of a plane
  static_assert(NI % 64 == 0, "interior lanes must fill whole wavefronts");
  static_assert(NR1 + 4 <= 64, "ring columns and corners must fit one wavefront");
};

struct Geo {
  int nx, ny, nz;
  long long n;
  int zchunk;
  int tiles_x, tiles_y, nchunks;
  int ntiles;  // tiles_x * tiles_y * nchunks
  // a z-slab of a grid decomposed along z: the two ghost planes below (zlo)
  // and above (zhi) come from the neighbouring slabs, [6][2][ny][nx] in the
  // storage type; null: the z rules of bound
  const void *zlo, *zhi;
};

template <int TY>
__device__ __forceinline__ int eidx(int x, int y) { return (x + 2) + TileShape<TY>::EW * (y + 2); }

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
    const float m2 = fm::fmad(q[3], q[3], fm::fmad(q[2], q[2], q[1] * q[1]));
    const float kin = 0.5f * m2 * rinv;
    const bool fast_flow = kin > float(P.kin_frac64) * q[4];
    float pr;
    if (P.dualenergy && kin * fm::rcp(q[4]) > float(P.de_thresh1)) {
      const double r64 = double(q[0]);
      const double k64 = 0.5 * (double(q[1]) * q[1] + double(q[2]) * q[2] + double(q[3]) * q[3]) / r64;
      // (the Newton form of the power: this rare branch must not raise the
      // register count of the whole FP32 kernel)
      const double p64 = double(q[5]) * fast::pow_g1<false>(P, r64);
      q[4] = float(k64 + p64 * P.g1inv);
      pr = float(p64);
    } else {
      if (fast_flow) {
        const double k64 = 0.5 * (double(q[1]) * q[1] + double(q[2]) * q[2] + double(q[3]) * q[3]) / double(q[0]);
        pr = float(P.g1 * (double(q[4]) - k64));
      } else {
        pr = float(P.g1) * (q[4] - kin);
      }
      if (P.dualenergy) q[5] = pr * fast::rpow_g1(P, q[0]);
    }
    w[4] = pr;
    return (pr <= 0.f) ? kErrNegPressure : kOk;
  }
};

// Arithmetic of the march.  Phys<false> is the fast form (physics_fast.h:
// refined hardware reciprocals and roots, cube-root powers, FMA contraction
// in the Riemann solvers; in FP64 the primitive recovery and the predictor
// keep the reference's roundings, with IEEE quotients from shared
// reciprocals).
// Phys<true> is the reference's own arithmetic (physics.h: IEEE division
// and square root, the reference's operation order; its translation unit is
// compiled without FMA contraction), so that every face state, flux and
// update is the one the reference forms, bit for bit except where the
// entropy takes a power (device pow against the host libm).
template <bool EXACT>
struct Phys;

template <>
struct Phys<false> {
  template <typename T>
  __device__ static __forceinline__ int q2w(const KParams &P, T q[6], T w[5], T &rinv) {
    return Prim<T>::q2w(P, q, w, rinv);
  }
  template <int LIMITER, typename T>
  __device__ static __forceinline__ T slope3(T wl, T wc, T wr) {
    return fast::slope3<LIMITER>(wl, wc, wr);
  }
  template <int D, typename T>
  __device__ static __forceinline__ void predictor_dir(const KParams &P, const T w[5], T rinv, const T s[5], T a[5]) {
    fast::predictor_dir<D>(P, w, rinv, s, a);
  }
  template <int SOLVER, int DIR, typename T>
  __device__ static __forceinline__ int riemann(const KParams &P, const T wl[5], const T wr[5], T f[6]) {
    return fast::riemann<SOLVER, DIR>(P, wl, wr, f);
  }
  template <typename T>
  __device__ static __forceinline__ double cfl_cell(const KParams &P, const T w[5], T rinv) {
    return fast::cfl_cell(P, w, rinv);
  }
  // temperature p / rho * temp0 (for thermal conduction)
  template <typename T>
  __device__ static __forceinline__ T temp(const KParams &P, const T w[5], T rinv) {
    return w[4] * rinv * T(P.temp0);
  }
};

// The reference's HLL and HLLC (riemann.f90:47-153) with its expressions
// and operation order (physics.h), written branch-free like the fast forms:
// the side is selected first and one flux / state / star state is formed.
// Every quotient is the IEEE one, formed as fm::div_cr from the refined
// reciprocal of its divisor (several quotients by one density share it),
// and every square root as fm::sqrt_pos: the sequences the compiler emits
// for IEEE division and square root, without their range scaling, so the
// results are the IEEE ones for the normal operands of the hydro step.  The
// entropy components take p / rho**g1 from the fast power (fastmath.h); they
// are the only ones a power enters, and the device and host powers differ
// in the last bits in any case.
namespace exact_bf {
template <int DIR>
__device__ __forceinline__ void flux(const KParams &P, const double w[5], double sq, double f[6]) {
  const double u2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
  const double ud = w[DIR];
  f[0] = w[0] * ud;
#pragma unroll
  for (int i = 1; i <= 3; i++) f[i] = (i == DIR) ? (w[0] * (ud * ud) + w[4]) : (w[0] * ud * w[i]);
  f[4] = ud * (0.5 * w[0] * u2 + P.gg1 * w[4]);
  f[5] = ud * sq;
}
// state (riemann.f90:207-239); p/g1 from the correctly rounded 1/g1 (g1inv)
__device__ __forceinline__ void state(const KParams &P, const double w[5], double sq, double q[6]) {
  q[0] = w[0];
  q[1] = w[0] * w[1];
  q[2] = w[0] * w[2];
  q[3] = w[0] * w[3];
  const double u2 = (w[1] * w[1]) + (w[2] * w[2]) + (w[3] * w[3]);
  q[4] = 0.5 * w[0] * u2 + fm::div_cr(w[4], P.g1, P.g1inv);
  q[5] = sq;
}
// sqrt(g p / rho) with y the refined reciprocal of rho
__device__ __forceinline__ double sound(const KParams &P, double rho, double p, double y) {
  return fm::sqrt_pos(fm::div_cr(P.g * p, rho, y));
}
template <int DIR>
__device__ __forceinline__ int hll(const KParams &P, const double wl[5], const double wr[5], double f[6]) {
  int code = (wl[4] <= 0.0 || wr[4] <= 0.0) ? kErrRiemannPressure : kOk;
  const double al = sound(P, wl[0], wl[4], fm::rcp(wl[0]));
  const double ar = sound(P, wr[0], wr[4], fm::rcp(wr[0]));
  const double sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const double sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  const bool left = 0.0 < sl, right = 0.0 > sr, mid = sl <= 0.0 && 0.0 <= sr;
  if (!left && !right && !mid && !code) code = kErrHll;
  const double sql = wl[4] * fast::rpow_g1(P, wl[0]);
  const double sqr = wr[4] * fast::rpow_g1(P, wr[0]);
  double fl[6], fr[6], ql[6], qr[6];
  flux<DIR>(P, wl, sql, fl);
  flux<DIR>(P, wr, sqr, fr);
  state(P, wl, sql, ql);
  state(P, wr, sqr, qr);
  const double den = sr - sl;
  const double yd = fm::rcp(den);
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const double fm_ = fm::div_cr(sr * fl[i] - sl * fr[i] + sl * sr * (qr[i] - ql[i]), den, yd);
    f[i] = left ? fl[i] : (right ? fr[i] : fm_);
  }
  return code;
}
template <int DIR>
__device__ __forceinline__ int hllc(const KParams &P, const double wl[5], const double wr[5], double f[6]) {
  int code = (wl[4] <= 0.0 || wr[4] <= 0.0) ? kErrRiemannPressure : kOk;
  const double yl = fm::rcp(wl[0]), yr = fm::rcp(wr[0]);
  const double al = sound(P, wl[0], wl[4], yl);
  const double ar = sound(P, wr[0], wr[4], yr);
  const double sl = dmin(wl[DIR] - al, wr[DIR] - ar);
  const double sr = dmax(wl[DIR] + al, wr[DIR] + ar);
  const double facl = wl[0] * (sl - wl[DIR]);
  const double facr = wr[0] * (sr - wr[DIR]);
  const double fd = facl - facr;
  const double sm = fm::div_cr(wr[4] - wl[4] + facl * wl[DIR] - facr * wr[DIR], fd, fm::rcp(fd));
  const int br = (0.0 < sl) ? 0
               : (0.0 > sr) ? 1
               : (sl < 0.0 && 0.0 < sm) ? 2
               : (sm <= 0.0 && 0.0 < sr) ? 3 : -1;
  if (br < 0 && !code) code = kErrHllc;
  const bool right = br == 1 || br == 3;
  double w[5];
#pragma unroll
  for (int i = 0; i < 5; i++) w[i] = right ? wr[i] : wl[i];
  const double y = right ? yr : yl;
  const double sq = w[4] * fast::rpow_g1(P, w[0]);
  flux<DIR>(P, w, sq, f);
  // star state (statex, riemann.f90:241-288) and f + s (qx - q)
  const double s = right ? sr : sl;
  const double ud = w[DIR];
  double q[6], qx[6];
  state(P, w, sq, q);
  const double ssm = s - sm;
  qx[0] = fm::div_cr(w[0] * (s - ud), ssm, fm::rcp(ssm));
#pragma unroll
  for (int i = 1; i <= 3; i++) qx[i] = ((i == DIR) ? sm : w[i]) * qx[0];
  const double energy = q[4];  // 0.5 rho u2 + p/g1, formed as in state
  const double rsu = w[0] * (s - ud);
  qx[4] = qx[0] * (fm::div_cr(energy, w[0], y) + (sm - ud) * (sm + fm::div_cr(w[4], rsu, fm::rcp(rsu))));
  qx[5] = qx[0] * fm::div_cr(sq, w[0], y);
  const bool star = br >= 2;
#pragma unroll
  for (int i = 0; i < 6; i++) f[i] = star ? f[i] + s * (qx[i] - q[i]) : f[i];
  return code;
}
}  // namespace exact_bf

template <>
struct Phys<true> {
  // q2w + pressure (comp.f90:26-85): the fast form's FP64 primitive already
  // forms every value as the reference does (IEEE quotients by rho from one
  // reciprocal), and returns the correctly rounded 1/rho; the dual-energy
  // branches take their powers from the fast form
  __device__ static __forceinline__ int q2w(const KParams &P, double q[6], double w[5], double &rinv) {
    return fast::q2w(P, q, w, rinv);
  }
  template <int LIMITER>
  __device__ static __forceinline__ double slope3(double wl, double wc, double wr) {
    return evh::slope3<LIMITER>(wl, wc, wr);
  }
  // muscl.f90:100-104: s(5)/rho is the IEEE quotient from rinv
  template <int D>
  __device__ static __forceinline__ void predictor_dir(const KParams &P, const double w[5], double rinv,
                                                       const double s[5], double a[5]) {
    fast::predictor_dir<D>(P, w, rinv, s, a);
  }
  template <int SOLVER, int DIR>
  __device__ static __forceinline__ int riemann(const KParams &P, const double wl[5], const double wr[5], double f[6]) {
    if (SOLVER == 1) return exact_bf::hll<DIR>(P, wl, wr, f);
    return exact_bf::hllc<DIR>(P, wl, wr, f);
  }
  // timestep.f90:39-44 for one cell
  __device__ static __forceinline__ double cfl_cell(const KParams &P, const double w[5], double rinv) {
    const double a = exact_bf::sound(P, w[0], w[4], rinv);
    const double sx = fabs(w[1]) + a, sy = fabs(w[2]) + a, sz = fabs(w[3]) + a;
    const double vx = fm::div_cr(P.dx, sx, fm::rcp(sx));
    const double vy = fm::div_cr(P.dy, sy, fm::rcp(sy));
    const double vz = fm::div_cr(P.dz, sz, fm::rcp(sz));
    return dmin(dmin(vx, vy), vz);
  }
  __device__ static __forceinline__ double temp(const KParams &P, const double w[5], double rinv) {
    return fm::div_cr(w[4], w[0], rinv) * P.temp0;
  }
};

// Ghost-mapped offsets (the index rules of `bound`) of column (gx, gy) and
// of plane gz, clamped into the ghost layers: partial tiles in x and y, and
// in z the prefetch of plane nz + 2 after the last plane (never used; it
// would fall outside the grid for nz = 2).  A column's offset is fixed for
// the whole march, so only the plane's is formed per plane.
__device__ __forceinline__ long long col_offset(const KParams &P, const Geo &G, int gx, int gy) {
  const bool per = P.boundtype == 2;
  const int cx = gx < -2 ? -2 : (gx > G.nx + 1 ? G.nx + 1 : gx);
  const int cy = gy < -2 ? -2 : (gy > G.ny + 1 ? G.ny + 1 : gy);
  return (long long)ghost_src(cx, G.nx, per) + (long long)G.nx * ghost_src(cy, G.ny, per);
}
__device__ __forceinline__ long long plane_offset(const KParams &P, const Geo &G, int gz) {
  const int cz = gz < -2 ? -2 : (gz > G.nz + 1 ? G.nz + 1 : gz);
  return (long long)G.nx * G.ny * ghost_src(cz, G.nz, P.boundtype == 2);
}

// raw q of the cell at offset c
template <typename S>
__device__ __forceinline__ void fetch_at(const Geo &G, const S *__restrict__ q, long long c, S raw[6]) {
#pragma unroll
  for (int m = 0; m < 6; m++) raw[m] = q[m * G.n + c];
}

// raw q of column offset c in plane gz (clamped as plane_offset does).
// Inside the grid, and on a side without a neighbour, that is q with the
// ghost rules of bound; in a z-slab (SLAB) the ghost planes beyond the slab
// are the neighbour's planes in the halo buffer of that side, [6][2][ny][nx]
// (the planes communicate receives, parallel.f90:317-381).  Whole grids take
// the kernels without the test.
template <bool SLAB, typename S>
__device__ __forceinline__ void fetch_plane(const KParams &P, const Geo &G, const S *__restrict__ q, int gz,
                                            long long c, S raw[6]) {
  if constexpr (SLAB) {
    const long long pl = (long long)G.nx * G.ny;
    const S *h = nullptr;
    if (gz < 0 && G.zlo) h = (const S *)G.zlo + pl * (gz < -2 ? 0 : gz + 2);
    else if (gz >= G.nz && G.zhi) h = (const S *)G.zhi + pl * (gz > G.nz ? 1 : 0);
    if (h) {
#pragma unroll
      for (int m = 0; m < 6; m++) raw[m] = h[m * 2 * pl + c];
      return;
    }
  }
  fetch_at(G, q, c + plane_offset(P, G, gz), raw);
}

// raw q at (gx, gy, gz), ghost-mapped
template <bool SLAB, typename S>
__device__ __forceinline__ void fetch_raw(const KParams &P, const Geo &G, const S *__restrict__ q, int gx, int gy,
                                          int gz, S raw[6]) {
  fetch_plane<SLAB>(P, G, q, gz, col_offset(P, G, gx, gy), raw);
}

// primitive of a fetched cell at column gx (for the reflecting wall)
template <bool EXACT, typename S, typename T>
__device__ __forceinline__ void prim_cell(const KParams &P, const Geo &G, const S raw[6], int gx, T qc[6], T wc[5],
                                          T &rinv) {
#pragma unroll
  for (int m = 0; m < 6; m++) qc[m] = T(raw[m]);
  Phys<EXACT>::q2w(P, qc, wc, rinv);
  if (P.boundtype == 3 && gx >= G.nx) {  // reflecting upper x (boundary.f90:57-64)
    wc[1] = -wc[1];
    wc[2] = -wc[2];
    wc[3] = -wc[3];
  }
}

// load q at (gx, gy, gz) (ghost-mapped) and recover w
template <bool EXACT, bool SLAB, typename S, typename T>
__device__ __forceinline__ void load_cell(const KParams &P, const Geo &G, const S *__restrict__ q, int gx, int gy,
                                          int gz, T qc[6], T wc[5], T &rinv) {
  S raw[6];
  fetch_raw<SLAB>(P, G, q, gx, gy, gz, raw);
  prim_cell<EXACT>(P, G, raw, gx, qc, wc, rinv);
}

template <typename T, int EA>
__device__ __forceinline__ void store_w(T (*W)[EA], int slot, const T w[5]) {
#pragma unroll
  for (int m = 0; m < 5; m++) W[m][slot] = w[m];
}

// LDS of one block
template <typename T, int TY>
struct Smem {
  T W[5][TileShape<TY>::EA];    // w of plane k, both halo rings
  T XB[10][TileShape<TY>::NFX]; // x face (fi, j), fi = 0..TX: [0,5) left state, [5,10) right state; then flux
  T YB[10][TileShape<TY>::NFY]; // y face (i, fj), fj = 0..TY
  T CR[17][TileShape<TY>::NI];  // per interior column: partial update [0,6), z flux below [6,12), z state below [12,17)
};

// One tile `lin` (x fastest, then y, then z chunk) of the box whose q starts
// at element qoff of qin_all/qout_all.  vmin_dst receives the CFL minimum of
// the tile's new state; errors carry the box index above bit 40 of the cell.
// With tout (single box only), the temperature of each new cell is stored
// there too, for the conduction pass that follows the sweep.  SLAB: the grid
// is a z-slab whose ghost planes may come from halo buffers (Geo::zlo, zhi).
template <typename T, typename S, int SOLVER, int LIMITER, int TY, bool EXACT, bool SLAB = false>
__device__ __forceinline__ void march(const KParams P, const Geo G, const S *__restrict__ qin_all,
                                      S *__restrict__ qout_all, long long qoff, int box, double dt, Status *st,
                                      unsigned long long *vmin_dst, int lin, Smem<T, TY> &sm,
                                      S *__restrict__ tout = nullptr) {
  using TS = TileShape<TY>;
  constexpr int NI = TS::NI, NR1 = TS::NR1, EW = TS::EW, EA = TS::EA;
  T(&W)[5][EA] = sm.W;
  T(&XB)[10][TS::NFX] = sm.XB;
  T(&YB)[10][TS::NFY] = sm.YB;
  T(&CR)[17][NI] = sm.CR;
  const int tid = threadIdx.x;
  const bool ring = tid >= NI;
  const int tile = lin % (G.tiles_x * G.tiles_y);
  const int chunk = lin / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * TX, y0 = (tile / G.tiles_x) * TY;
  const int z0 = chunk * G.zchunk;
  const int z1 = min(z0 + G.zchunk, G.nz);
  const S *__restrict__ qin = qin_all + qoff;
  S *__restrict__ qout = qout_all + qoff;
  const unsigned long long tag = (unsigned long long)box << 40;  // box index above the cell index

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
  const int me = eidx<TY>(cx, cy);

  const T dt2 = T(-0.5 * dt);
  const T dtx = T(dt * P.dxinv), dty = T(dt * P.dyinv), dtz = T(dt * P.dzinv);
  const T half = T(0.5);

  // qc, qn: the (dual-energy corrected) q of planes k and k+1, the base of
  // the update, kept in registers so that q is read from HBM once
  T wm[5], wc[5], wp[5], qn[6], qc[6], tq[6];
  T rm, rc, rp;
  // The CFL minimum min_c 1 / max_d (|u_d| + a) / dx_d.  The FP64 fast
  // form reduces -max_d (..) per cell and inverts once per lane; a lane
  // without cells gives 1e6 either way.
  constexpr bool kSpeedKey = !EXACT && sizeof(T) == 8;
  double vmin = kSpeedKey ? -1e-6 : 1e6;
  // first Riemann error and first negative pressure of the new state seen by
  // this lane (code, plane); reported once after the march
  int err = 0, err_k = 0, nerr = 0, nerr_k = 0;
  // q^{n+1} of plane k-1, stored in S2; its primitive, CFL limit and
  // temperature are formed after the x/y solves (S4), where the interior
  // waves would otherwise wait for the ring wave
  T o[6];
  bool pend = false;
  long long oc = 0;

  // global loads run one plane ahead: pre holds the raw q of this column's
  // next plane and pre2 the ring wave's second-ring cell, so that no wave
  // stalls on memory right after issuing a load
  S pre[6], pre2[6];
  const long long own_xy = col_offset(P, G, gx, gy);
  const long long r2_xy = col_offset(P, G, x0 + r2x, y0 + r2y);
  // the ring wave's second-ring cell: its w of the current plane (rw2) is
  // formed one plane ahead, in the ring wave's idle time of S2
  T rw2[5];
#pragma unroll
  for (int m = 0; m < 6; m++) { qc[m] = T(0); qn[m] = T(0); }
  load_cell<EXACT, SLAB, S, T>(P, G, qin, gx, gy, z0 - 2, tq, wm, rm);
  load_cell<EXACT, SLAB, S, T>(P, G, qin, gx, gy, z0 - 1, tq, wc, rc);
  fetch_plane<SLAB>(P, G, qin, z0, own_xy, pre);
  {
    T rr2;
    fetch_plane<SLAB>(P, G, qin, z0 - 1, r2_xy, pre2);
    prim_cell<EXACT>(P, G, pre2, x0 + r2x, tq, rw2, rr2);
    fetch_plane<SLAB>(P, G, qin, z0, r2_xy, pre2);
  }

  for (int k = z0 - 1; k <= z1; k++) {
    const bool out_plane = k >= z0 && k < z1;
    // ---- S1: primitive of plane k+1 (own column); rings of plane k to LDS.
    // Every lane forms and fetches (the ring wave's corner and idle lanes on
    // clamped coordinates), so that no branch makes a load wait for the
    // fetches issued after it.
    prim_cell<EXACT>(P, G, pre, gx, qn, wp, rp);
    fetch_plane<SLAB>(P, G, qin, k + 2, own_xy, pre);
    if (ring && r2x > -3) store_w<T, EA>(W, eidx<TY>(r2x, r2y), rw2);
    if (owns_col) store_w<T, EA>(W, me, wc);
    __syncthreads();

    // ---- S2: slopes, predictor and face states of plane k
    T sx[5], sy[5], sz[5], ev[5];
    if (owns_col) {
#pragma unroll
    for (int m = 0; m < 5; m++) {
      sx[m] = Phys<EXACT>::template slope3<LIMITER>(W[m][me - 1], wc[m], W[m][me + 1]);
      sy[m] = Phys<EXACT>::template slope3<LIMITER>(W[m][me - EW], wc[m], W[m][me + EW]);
      sz[m] = Phys<EXACT>::template slope3<LIMITER>(wm[m], wc[m], wp[m]);
    }
    if (P.secondorder) {
      T a1[5], a2[5], a3[5];
      Phys<EXACT>::template predictor_dir<0>(P, wc, rc, sx, a1);
      Phys<EXACT>::template predictor_dir<1>(P, wc, rc, sy, a2);
      Phys<EXACT>::template predictor_dir<2>(P, wc, rc, sz, a3);
      fast::evo5(P, dt2, a1, a2, a3, ev);  // muscl.f90:106-107
    } else {
#pragma unroll
      for (int m = 0; m < 5; m++) { ev[m] = T(0); sx[m] = T(0); sy[m] = T(0); sz[m] = T(0); }
    }
    // face states (muscl.f90:207-210) in the reference's order:
    // (w + 0.5 s) + evo on the high side, (w - 0.5 s) + evo on the low side
    if (!ring) {
      const int fxl = cx + (TX + 1) * cy, fyb = cx + TX * cy;
#pragma unroll
      for (int m = 0; m < 5; m++) {
        XB[5 + m][fxl] = fm::fmad(-half, sx[m], wc[m]) + ev[m];       // right state of the left face
        XB[m][fxl + 1] = fm::fmad(half, sx[m], wc[m]) + ev[m];       // left state of the right face
        YB[5 + m][fyb] = fm::fmad(-half, sy[m], wc[m]) + ev[m];
        YB[m][fyb + TX] = fm::fmad(half, sy[m], wc[m]) + ev[m];
      }
    } else if (side == 0) {
#pragma unroll
      for (int m = 0; m < 5; m++) XB[m][(TX + 1) * cy] = fm::fmad(half, sx[m], wc[m]) + ev[m];
    } else if (side == 1) {
#pragma unroll
      for (int m = 0; m < 5; m++) XB[5 + m][TX + (TX + 1) * cy] = fm::fmad(-half, sx[m], wc[m]) + ev[m];
    } else if (side == 2) {
#pragma unroll
      for (int m = 0; m < 5; m++) YB[m][cx] = fm::fmad(half, sy[m], wc[m]) + ev[m];
    } else {
#pragma unroll
      for (int m = 0; m < 5; m++) YB[5 + m][cx + TX * TY] = fm::fmad(-half, sy[m], wc[m]) + ev[m];
    }
    // z face k-1/2 (interior lanes) and the end of plane k-1; the
    // scheduling barrier keeps this solve apart from the x/y solves
    __builtin_amdgcn_sched_barrier(0);
    if (!ring) {
      T zl[5], zr[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        zl[m] = fm::fmad(-half, sz[m], wc[m]) + ev[m];
        zr[m] = fm::fmad(half, sz[m], wc[m]) + ev[m];
      }
      if (k >= z0) {
        T zrp[5], fz[6];
#pragma unroll
        for (int m = 0; m < 5; m++) zrp[m] = CR[12 + m][tid];
        const int e = Phys<EXACT>::template riemann<SOLVER, 3>(P, zrp, zl, fz);
        if (e && inside && !err) { err = e; err_k = k; }
        if (k - 1 >= z0 && inside) {
#pragma unroll
          for (int m = 0; m < 6; m++) o[m] = fm::fmad(dtz, CR[6 + m][tid] - fz[m], CR[m][tid]);
          oc = (long long)gx + (long long)G.nx * ((long long)gy + (long long)G.ny * (k - 1));
#pragma unroll
          for (int m = 0; m < 6; m++) qout[m * G.n + oc] = S(o[m]);
          pend = true;
        }
#pragma unroll
        for (int m = 0; m < 6; m++) CR[6 + m][tid] = fz[m];
      }
#pragma unroll
      for (int m = 0; m < 5; m++) CR[12 + m][tid] = zr[m];
    }
    }  // owns_col
    if (ring) {  // second-ring w of plane k+1, stored in S1 of plane k+1
      T rr2;
      prim_cell<EXACT>(P, G, pre2, x0 + r2x, tq, rw2, rr2);
      fetch_plane<SLAB>(P, G, qin, k + 2, r2_xy, pre2);
    }
    __syncthreads();

    // ---- S3: x and y faces of plane k
    if (out_plane) {
      if (!ring) {
        // interior lane (i,j): its left x face and bottom y face, as one
        // straight-line block so that the two solves interleave
        const int fxi = cx + (TX + 1) * cy, fyi = cx + TX * cy;
        T xl[5], xr[5], fx[6], yl[5], yr[5], fy[6];
#pragma unroll
        for (int m = 0; m < 5; m++) {
          xl[m] = XB[m][fxi];
          xr[m] = XB[5 + m][fxi];
          yl[m] = YB[m][fyi];
          yr[m] = YB[5 + m][fyi];
        }
        const int ex = Phys<EXACT>::template riemann<SOLVER, 1>(P, xl, xr, fx);
        const int ey = Phys<EXACT>::template riemann<SOLVER, 2>(P, yl, yr, fy);
        // faces of cells outside the box (partial tiles) do not count
        const int e = (gy < G.ny && gx <= G.nx && ex) ? ex : ((gx < G.nx && gy <= G.ny) ? ey : 0);
        const bool take = e && !err;
        err = take ? e : err;
        err_k = take ? k : err_k;
#pragma unroll
        for (int m = 0; m < 6; m++) {
          XB[m][fxi] = fx[m];
          YB[m][fyi] = fy[m];
        }
      } else {
        // ring wave: the TY right-edge x faces (lanes 0..TY-1) and the 16
        // top-edge y faces (lanes TY..TY+15) in one solve, so that the wave
        // does not run two: a y face is the x face of the states with u and
        // v exchanged.  The solvers are bit-symmetric under that exchange:
        // |u|^2 is formed symmetrically and every other term by its role
        // (the FP32 build has no contraction).
        if (lane < TY + TX) {
          const bool isy = lane >= TY;
          T *fb = isy ? &YB[0][(lane - TY) + TX * TY] : &XB[0][TX + (TX + 1) * lane];
          const int st_ = isy ? TS::NFY : TS::NFX;
          T wl[5], wr[5], f[6];
#pragma unroll
          for (int m = 0; m < 5; m++) { wl[m] = fb[m * st_]; wr[m] = fb[(5 + m) * st_]; }
          const T l1 = wl[1], r1 = wr[1];
          wl[1] = isy ? wl[2] : wl[1];
          wl[2] = isy ? l1 : wl[2];
          wr[1] = isy ? wr[2] : wr[1];
          wr[2] = isy ? r1 : wr[2];
          const int e = Phys<EXACT>::template riemann<SOLVER, 1>(P, wl, wr, f);
          const T f1 = f[1];
          f[1] = isy ? f[2] : f[1];
          f[2] = isy ? f1 : f[2];
          const bool counts = isy ? (x0 + lane - TY < G.nx && y0 + TY <= G.ny) : (x0 + TX <= G.nx && y0 + lane < G.ny);
          if (e && !err && counts) { err = e; err_k = k; }
#pragma unroll
          for (int m = 0; m < 6; m++) fb[m * st_] = f[m];
        }
      }
    }
    __syncthreads();

    // ---- S4: x/y differences of plane k
    if (!ring && out_plane) {
      const int fxl = cx + (TX + 1) * cy, fyb = cx + TX * cy;
      T part[6];
#pragma unroll
      for (int m = 0; m < 6; m++) part[m] = fm::fmad(dtx, XB[m][fxl] - XB[m][fxl + 1], qc[m]);
#pragma unroll
      for (int m = 0; m < 6; m++) CR[m][tid] = fm::fmad(dty, YB[m][fyb] - YB[m][fyb + TX], part[m]);
    }
    // the end of plane k-1: primitive of the new state for the CFL minimum
    // of the next step, its first negative pressure, its temperature
    if (pend) {
      T on[6], wn[5], rn;
#pragma unroll
      for (int m = 0; m < 6; m++) on[m] = T(S(o[m]));
      const int ep = Phys<EXACT>::q2w(P, on, wn, rn);
      if (ep && !nerr) { nerr = ep; nerr_k = k - 1; }
      if constexpr (kSpeedKey) vmin = dmin(vmin, -fast::cfl_speed(P, wn, rn));
      else vmin = dmin(vmin, Phys<EXACT>::cfl_cell(P, wn, rn));
      if (tout) tout[oc] = S(Phys<EXACT>::temp(P, wn, rn));
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

  if (err || nerr) {  // the lane's own column (clamped into the box) locates the error
    const long long ex = min(max(gx, 0), G.nx - 1), ey = min(max(gy, 0), G.ny - 1);
    if (err) report(st, err, tag | (unsigned long long)(ex + G.nx * (ey + (long long)G.ny * err_k)));
    if (nerr) report_next(st, nerr, tag | (unsigned long long)(ex + G.nx * (ey + (long long)G.ny * nerr_k)));
  }
  if constexpr (kSpeedKey) vmin = fm::rcp(-vmin);
  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((tid & 63) == 0) atomic_min_pos(vmin_dst, vmin);
}


}  // namespace core
}  // namespace evh
