// gravity.hip -- self-gravity on the device: the Poisson solve of
// gravpotential (gravitation.f90:26-88) and gravintegration with grav_force
// (gravitation.f90:172-254), one rank.
//
// Potential.  The source grav (rho - mean rho) (or, supercomoving,
// cosmo grav (rho / fb - 1)) goes through a real-to-complex FFT (FP64 in
// both precision forms), is multiplied by the Green's function of
// multigreen, -h2 / (sin^2(kfx2 kx) + sin^2(kfy2 ky) + sin^2(kfz2 kz)) (0 at
// k = 0), and comes back through the complex-to-real FFT.  The reference
// transforms complex to complex and keeps the real part; since its source is
// real, that real part is the transform of the Hermitian half of the product,
// which is what the real transforms form once the Green's function is taken
// as (G(k) + G(-k)) / 2.  The two terms are equal bit for bit except for odd
// grid sizes, where kvalue (comp.f90:241-252) gives the bins next to the
// middle frequencies whose sines differ by the rounding of the
// single-precision pi of global.f90:98.  The potential stays in FP64 in the
// mixed form too: the force is a difference of neighbouring potentials.
// The mean density only sets the (0,0,0) mode, which the Green's function
// zeroes; on power-of-two grids the x pass sums rho as it reads it and the
// next call takes its mean from those sums (an upload invalidates them).
//
// Transforms.  When nz is a power of two, the z transform is the pencil
// kernel below: a block takes a tile of consecutive kx columns of one y row,
// transforms each column along z with Stockham stages of radix 8 (the first
// reading the spectrum, the others the tile in LDS), applies the Green's
// function to the last forward stage's outputs, transforms back and writes
// the tile from the last inverse stage, so that the spectrum is read and
// written once between the forward and inverse x-y transforms.  When nx and
// ny are powers of two as well, the x and y transforms are this file's too:
// k_xfwd forms the source of two real x rows as it loads them, transforms
// them as one complex row and splits the result into the rows' half
// spectra; y pencils transform forward, the z pass follows, y pencils
// transform back, and k_xinv joins two half spectra into one complex row
// whose inverse transform is the two rows of phi.  No pass of its own forms
// the source, and the spectrum is read and written once per pass.  With
// other nx or ny, hipFFT's 2-D x-y transforms (one per z plane) surround the
// z pass; other nz take hipFFT's 3-D plans and k_green.  2-D (nz = 1) and
// 1-D (ny = nz = 1) grids take hipFFT's 2-D and 1-D plans and k_green: the
// reference's 3-D transform is the identity along a one-cell axis, whose
// only wavenumber is 0 (sin^2 0 in the Green's function), and its force
// component is 0 (gravitation.f90:209,220,232).
//
// Force and update.  k_gravint marches z planes of 64 x 4 column tiles with
// the potential of the tile and its one-cell ring in LDS (periodic wrap, as
// grav_force), adds dt q(2:4).f to q(5) and dt q(1) f to q(2:4) in place,
// and feeds the primitive of the new state into the CFL minimum of the next
// hydrostep (and, with conduction on, stores T for the conduction pass).
//
// This unit is compiled without FMA contraction: the update and the Green's
// function follow the reference's expressions and order.
#include "evh_internal.h"
#include "physics_fast.h"

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

namespace evh {
namespace {

constexpr int kSumBlocks = 1024;
constexpr int GX = 64, GY = 4, GT = GX * GY;  // force tile and threads
constexpr int GRX = GX + 2, GRY = GY + 2;
constexpr int GRING = 2 * GX + 2 * GY;
constexpr int kZChunk = 128;

struct GParams {
  double grav, fb, h2, n3, n3inv;
  double dx2inv, dy2inv, dz2inv;
  int supercomoving;
};

struct GGeo {
  int nx, ny, nz;
  long long n;
  int zchunk, tiles_x, tiles_y, ntiles;
};

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ---- mean density: Kahan per thread, pairwise tree per block, then one block
template <typename S>
__global__ void __launch_bounds__(256) k_rho_partials(const S *__restrict__ q, long long n, double *part) {
  __shared__ double red[256];
  // four compensated sums over four interleaved streams, so that four loads
  // are in flight per thread; then their compensated total
  double s[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
  const long long g = (long long)gridDim.x * 256;
  long long i = blockIdx.x * 256LL + threadIdx.x;
  for (; i + 3 * g < n; i += 4 * g) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = double(q[i + u * g]);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double y = v[u] - c[u];
      const double t = s[u] + y;
      c[u] = (t - s[u]) - y;
      s[u] = t;
    }
  }
  for (; i < n; i += g) {
    const double y = double(q[i]) - c[0];
    const double t = s[0] + y;
    c[0] = (t - s[0]) - y;
    s[0] = t;
  }
  double tot = 0.0, ct = 0.0;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const double y = (s[u] - c[u]) - ct;
    const double t = tot + y;
    ct = (t - tot) - y;
    tot = t;
  }
  red[threadIdx.x] = tot;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kSumBlocks) k_rho_mean(double *part, double n3) {
  __shared__ double red[kSumBlocks];
  red[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int w = kSumBlocks / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[kSumBlocks] = red[0] / n3;  // meanrho = sum / n3
}

// the mean density from the block totals of the previous call's x pass
// (Neumaier sums per thread, then a pairwise tree)
__global__ void __launch_bounds__(kSumBlocks) k_rho_mean_from(const double *__restrict__ sums, int nsums, double n3,
                                                            double *part) {
  __shared__ double red[kSumBlocks];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nsums; i += kSumBlocks) {
    const double v = sums[i], t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  red[threadIdx.x] = s + c;
  __syncthreads();
  for (int w = kSumBlocks / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[kSumBlocks] = red[0] / n3;
}

// ---- the source of gravitation.f90:33-42
template <typename S>
__global__ void __launch_bounds__(256) k_grav_source(const S *__restrict__ q, long long n, double *__restrict__ phi,
                                                     GParams G, double cosmo_grav, const double *__restrict__ mean) {
  const double m = G.supercomoving ? 0.0 : mean[kSumBlocks];
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double rho = double(q[i]);
    phi[i] = G.supercomoving ? cosmo_grav * (rho / G.fb - 1.0) : G.grav * (rho - m);
  }
}

// ---- z pencils of the spectrum (power-of-two nz)
#ifndef EVH_PENCIL_LOG
#define EVH_PENCIL_LOG 12
#endif
constexpr int kPencilLog = EVH_PENCIL_LOG;
constexpr int kPencilElems = 1 << kPencilLog;  // complex values per block, 16 B each in LDS
constexpr int kPencilThreads = kPencilElems / 8;  // 8 values per thread and stage
#ifndef EVH_PENCIL_WAVES
#define EVH_PENCIL_WAVES 4
#endif
constexpr int kPencilWaves = EVH_PENCIL_WAVES;  // waves per SIMD the register budget allows

struct Pencils {
  int ln;            // log2 of the transform length (nz or ny)
  int nxh;           // kx values of a row
  int ktiles;        // tiles of kx per outer index
  long long s_elem;  // stride of the transform index, z or y (complex values)
  long long s_outer; // stride of the outer index, y for the z transform, z for the y transform
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// The radix-R DFT of v in place, y_s = sum_r v_r w^(r s) with
// w = exp(SIGN 2 pi i / R): radix 4 as two radix-2 layers, radix 8 as two
// radix-4 DFTs of the even and odd terms joined by w^s.
template <int SIGN>
__device__ __forceinline__ double2 mul_i(double2 d) {  // SIGN i d
  return SIGN < 0 ? make_double2(d.y, -d.x) : make_double2(-d.y, d.x);
}

template <int SIGN, int R>
__device__ __forceinline__ void dft(double2 *v) {
  if (R == 2) {
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if (R == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 a2 = cadd(v[1], v[3]), a3 = mul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[1] = cadd(a1, a3);
    v[2] = csub(a0, a2);
    v[3] = csub(a1, a3);
  } else {
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft<SIGN, 4>(e);
    dft<SIGN, 4>(o);
    constexpr double c = 0.70710678118654752440;  // 1/sqrt(2)
    // w = c (1 + SIGN i), w^2 = SIGN i, w^3 = c (-1 + SIGN i)
    const double2 o1 = make_double2(c * (o[1].x - SIGN * o[1].y), c * (o[1].y + SIGN * o[1].x));
    const double2 o2 = mul_i<SIGN>(o[2]);
    const double2 o3 = make_double2(c * (-o[3].x - SIGN * o[3].y), c * (-o[3].y + SIGN * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
}

// where a stage reads and writes: the spectrum in HBM or the tile in LDS,
// and the last forward stage multiplies its outputs by the Green's function
// kFromSource / kToPhi: the x transform's first stage reads the source of
// two real rows of rho (one as the real, one as the imaginary part) and its
// inverse's last stage writes the two real rows of phi; kFromHalves: the
// inverse x transform's first stage forms that complex row from the two
// rows' half spectra.
enum { kFromLds, kFromSpec, kFromSource, kFromHalves };
enum { kToLds, kToLdsGreen, kToSpec, kToPhi };

struct PencilTile {
  double2 *X;     // tk columns of n values in LDS, X[i * tk + t]
  double2 *base;  // the tile's first column at index 0 of the transform axis
  int cols, kx0;
  int outer;      // y for the z transform, z for the y transform
  // x transforms: column t is the row pair (2t, 2t+1) after the tile's first
  // row; base is that row's half spectrum, rows s_elem values apart
  const void *rho;    // q(1) of the first row (storage type)
  double *phi;        // phi of the first row
  double mean, cosmo_grav;
  double *racc;       // k_xfwd: this thread's sum of the densities it reads (16 values at nx = 512)
};

// the source of gravpotential (gravitation.f90:26-53) of one density, as
// k_grav_source forms it; SC: supercomoving
template <bool SC>
__device__ __forceinline__ double grav_source(double rho, const GParams &G, double mean, double cosmo_grav) {
  return SC ? cosmo_grav * (rho / G.fb - 1.0) : G.grav * (rho - mean);
}

// multigreen (gravitation.f90:55-89) at (kx, ky, kz) = (i, j, k): the
// Green's function (symmetrised, see above) times the value divided by n3,
// as the complex quotient v / (n3 + 0i) (QUOTIENT) or as the product with
// 1/n3, which differs from it by an ulp
template <bool QUOTIENT>
__device__ __forceinline__ double2 green_apply(double2 v, int i, int j, int k, int nx, int ny, int nz,
                                               const double *__restrict__ sin2, const GParams &G) {
  const double *sx = sin2, *sy = sin2 + nx, *sz = sin2 + nx + ny;
  double green;
  if (i == 0 && j == 0 && k == 0) {
    green = 0.0;
  } else {
    const int im = i == 0 ? 0 : nx - i, jm = j == 0 ? 0 : ny - j, km = k == 0 ? 0 : nz - k;
    const double g1 = -G.h2 / (sx[i] + sy[j] + sz[k]);
    const double g2 = -G.h2 / (sx[im] + sy[jm] + sz[km]);
    green = g1 == g2 ? g1 : 0.5 * (g1 + g2);
  }
  double2 o;
  if (QUOTIENT) {
    const double den = G.n3 * G.n3 + 0.0 * 0.0;
    o.x = (v.x * G.n3 + v.y * 0.0) / den * green;
    o.y = (v.y * G.n3 - v.x * 0.0) / den * green;
  } else {
    const double f = green * G.n3inv;
    o.x = v.x * f;
    o.y = v.y * f;
  }
  return o;
}

// One Stockham stage of radix R after ns = 2^LNS on the tk = kPencilElems/n
// columns of the tile (n = 2^LN), tw[e] = exp(-2 pi i e / n), SIGN -1
// forward and +1 inverse (unnormalised).  Butterfly j < n/R of column t
// reads i = j + r n/R (r < R), twiddles it by w^(k r n / (ns R)) with
// k = j mod ns, and writes its R outputs to (j/ns) ns R + k + s ns; after the
// last stage that index is the frequency.  Every LDS read of a stage
// completes (barrier) before its LDS writes, and its writes before the next
// stage.
// The tile in LDS: columns (ROWS false) hold value i of column t at
// i tk + t, and the lanes of a butterfly index j run over t first; rows
// (ROWS true, the x transforms, whose rows are contiguous in HBM) hold it at
// t rp + i + i/8 (rp = n + n/8: one pad value per 8, so that the stride-R
// writes of a stage fall into different banks), and the lanes run over j
// first.
template <int LN, bool ROWS>
struct TileIndex {
  static constexpr int N = 1 << LN, LTK = kPencilLog - LN, TK = 1 << LTK;
  static constexpr int RP = N + N / 8;
  __device__ static __forceinline__ int at(int i, int t) { return ROWS ? t * RP + i + (i >> 3) : i * TK + t; }
};
template <bool ROWS>
constexpr int tile_elems() { return ROWS ? kPencilElems + kPencilElems / 8 : kPencilElems; }

template <int SIGN, int LN, int LNS, int R, int SRC, int DST, typename S, bool ROWS, bool SC>
__device__ __forceinline__ void pencil_stage(const PencilTile &pt, long long s_elem,
                                             const double2 *__restrict__ tw, const double *__restrict__ sin2,
                                             const GParams &G, int nx, int ny, int nz) {
  constexpr int LR = R == 8 ? 3 : R == 4 ? 2 : 1;
  constexpr int NB = kPencilElems / (R * kPencilThreads);  // butterflies per thread
  constexpr int LTK = kPencilLog - LN, TK = 1 << LTK;
  constexpr int NQ = 1 << (LN - LR), NS = 1 << LNS;
  constexpr int LSTEP = LN - LNS - LR;  // log2 of n / (ns R)
  using TI = TileIndex<LN, ROWS>;
  double2 v[NB][R];
  int dst[NB], tcol[NB];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int b = threadIdx.x + kPencilThreads * u;
    const int t = ROWS ? b >> (LN - LR) : b & (TK - 1);
    const int j = ROWS ? b & (NQ - 1) : b >> LTK;
    const int k = j & (NS - 1);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = j + r * NQ;
      double2 x;
      if (SRC == kFromSpec) {
        x = t < pt.cols ? pt.base[i * s_elem + t] : make_double2(0.0, 0.0);
      } else if (SRC == kFromSource) {
        x = make_double2(0.0, 0.0);
        if (t < pt.cols) {
          const S *r0 = (const S *)pt.rho + (long long)(2 * t) * (1 << LN);
          const double ra = double(r0[i]), rb = double(r0[i + (1 << LN)]);
          x.x = grav_source<SC>(ra, G, pt.mean, pt.cosmo_grav);
          x.y = grav_source<SC>(rb, G, pt.mean, pt.cosmo_grav);
          if constexpr (!SC) pt.racc[0] += ra + rb;  // the densities' sum, for the next call's mean
        }
      } else if (SRC == kFromHalves) {
        // Z(k) = A(k) + i B(k) over the whole row, A and B Hermitian; the
        // imaginary parts at k = 0 and n/2 are dropped, as a C2R transform does
        x = make_double2(0.0, 0.0);
        if (t < pt.cols) {
          constexpr int N = 1 << LN;
          const bool lo = i <= N / 2;
          const double2 *ra = pt.base + (long long)(2 * t) * s_elem + (lo ? i : N - i);
          double2 a = ra[0], b = ra[s_elem];
          if (!lo) { a.y = -a.y; b.y = -b.y; }
          if (i == 0 || i == N / 2) { a.y = 0.0; b.y = 0.0; }
          x = make_double2(a.x - b.y, a.y + b.x);
        }
      } else {
        x = pt.X[TI::at(i, t)];
      }
      if (r > 0 && LNS > 0) {  // the first stage's twiddles are 1
        double2 w = tw[(k * r) << LSTEP];
        if (SIGN > 0) w.y = -w.y;
        x = cmul(x, w);
      }
      v[u][r] = x;
    }
    dst[u] = ((j >> LNS) << (LNS + LR)) + k;
    tcol[u] = t;
  }
  // LDS read before LDS write, or (a one-stage transform) the spectrum read
  // by every lane before any lane overwrites it
  constexpr bool TO_LDS = DST == kToLds || DST == kToLdsGreen;
  if ((SRC == kFromLds && TO_LDS) || (SRC == kFromSpec && DST == kToSpec)) __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; u++) {
    dft<SIGN, R>(v[u]);
    const int t = tcol[u];
#pragma unroll
    for (int s = 0; s < R; s++) {
      const int i = dst[u] + s * NS;
      if (DST == kToSpec) {
        if (t < pt.cols) {
          double2 *o = pt.base + i * s_elem + t;  // streamed: no later read in this pass
          __builtin_nontemporal_store(v[u][s].x, &o->x);
          __builtin_nontemporal_store(v[u][s].y, &o->y);
        }
      } else if (DST == kToPhi) {
        if (t < pt.cols) {
          double *p0 = pt.phi + (long long)(2 * t) * (1 << LN);
          __builtin_nontemporal_store(v[u][s].x, p0 + i);
          __builtin_nontemporal_store(v[u][s].y, p0 + i + (1 << LN));
        }
      } else if (DST == kToLdsGreen) {
        pt.X[TI::at(i, t)] = t < pt.cols ? green_apply<false>(v[u][s], pt.kx0 + t, pt.outer, i, nx, ny, nz, sin2, G)
                                         : v[u][s];
      } else {
        pt.X[TI::at(i, t)] = v[u][s];
      }
    }
  }
  if (TO_LDS) __syncthreads();
}

// The stages of one transform from ns = 2^LNS on: radix 8 while three or
// more bits of log2 n remain, then one stage of radix 4 or 2.  With
// FROM_SPEC the first stage reads the spectrum (else the tile in LDS); the
// last stage writes LAST_DST.
template <int SIGN, int LN, int LNS, int FIRST_SRC, int LAST_DST, typename S = double, bool ROWS = false,
          bool SC = false>
__device__ __forceinline__ void pencil_stages(const PencilTile &pt, long long s_elem,
                                              const double2 *__restrict__ tw, const double *__restrict__ sin2,
                                              const GParams &G, int nx, int ny, int nz) {
  constexpr int REM = LN - LNS;
  constexpr int R = REM >= 3 ? 8 : REM == 2 ? 4 : 2;
  constexpr int LR = R == 8 ? 3 : R == 4 ? 2 : 1;
  constexpr bool LAST = LNS + LR == LN;
  constexpr int SRC = LNS == 0 ? FIRST_SRC : kFromLds;
  constexpr int DST = LAST ? LAST_DST : kToLds;
  pencil_stage<SIGN, LN, LNS, R, SRC, DST, S, ROWS, SC>(pt, s_elem, tw, sin2, G, nx, ny, nz);
  if constexpr (!LAST)
    pencil_stages<SIGN, LN, LNS + LR, FIRST_SRC, LAST_DST, S, ROWS, SC>(pt, s_elem, tw, sin2, G, nx, ny, nz);
}

// what a pencil launch does: the z transform with the Green's function, or
// the forward or the inverse y transform
enum { kPencilGreenZ, kPencilFwdY, kPencilInvY };

// A tile of tk consecutive kx columns at one outer index, n = 2^LN along
// the transform axis.  kPencilGreenZ: the forward transform along z, the
// Green's function on the last forward stage's outputs, and the inverse
// transform; kPencilFwdY / kPencilInvY: one transform along y.  The first
// stage reads the spectrum and the last writes it, so the tile passes
// through LDS between stages only.
template <int LN, int MODE>
__global__ void __launch_bounds__(kPencilThreads) __attribute__((amdgpu_waves_per_eu(kPencilWaves, kPencilWaves)))
k_pencils(double2 *__restrict__ spec, Pencils pc, const double2 *__restrict__ tw,
          const double *__restrict__ sin2, GParams G, int nx, int ny, int nz) {
  __shared__ double2 X[kPencilElems];
  PencilTile pt;
  pt.X = X;
  pt.outer = blockIdx.x / pc.ktiles;
  pt.kx0 = (blockIdx.x - pt.outer * pc.ktiles) << (kPencilLog - LN);
  pt.cols = min(1 << (kPencilLog - LN), pc.nxh - pt.kx0);
  pt.base = spec + pt.outer * pc.s_outer + pt.kx0;
  if (MODE == kPencilGreenZ) {
    pencil_stages<-1, LN, 0, kFromSpec, kToLdsGreen>(pt, pc.s_elem, tw, sin2, G, nx, ny, nz);
    pencil_stages<1, LN, 0, kFromLds, kToSpec>(pt, pc.s_elem, tw, sin2, G, nx, ny, nz);
  } else {
    pencil_stages<MODE == kPencilFwdY ? -1 : 1, LN, 0, kFromSpec, kToSpec>(pt, pc.s_elem, tw, sin2, G, nx, ny, nz);
  }
}

// The x transforms, n = nx = 2^LN, on tiles of tk row pairs (rows of nx
// values in (y, z) order; ny is even, so a pair shares its z plane).  Row
// pair t is transformed as the complex row z = a + i b, Z = A + i B.
// k_xfwd: a and b are the source of rows 2t and 2t+1 (formed from rho on
// the load); after the forward transform, A(k) = (Z(k) + conj Z(n-k)) / 2 and
// B(k) = (Z(k) - conj Z(n-k)) / 2i, k <= n/2, go to the two rows of the half
// spectrum.  This is the real-to-complex transform of each row, two rows per
// complex transform.  SC: the supercomoving source.
// Without SC the pass also sums the densities it reads, one total per block
// into rho_sum, from which the next call takes its mean.
template <typename S, int LN, bool SC>
__global__ void __launch_bounds__(kPencilThreads) __attribute__((amdgpu_waves_per_eu(kPencilWaves, kPencilWaves)))
k_xfwd(const S *__restrict__ rho, double2 *__restrict__ spec, long long pitch, long long npairs,
       const double2 *__restrict__ tw, GParams G, double cosmo_grav, const double *__restrict__ mean_part,
       double *__restrict__ rho_sum) {
  constexpr int N = 1 << LN, TK = 1 << (kPencilLog - LN), NH = N / 2 + 1;
  using TI = TileIndex<LN, true>;
  __shared__ double2 X[tile_elems<true>()];
  PencilTile pt;
  pt.X = X;
  const long long p0 = (long long)blockIdx.x * TK;  // first pair of the tile
  pt.cols = (int)min((long long)TK, npairs - p0);
  pt.kx0 = 0;
  pt.outer = 0;
  pt.base = spec + 2 * p0 * pitch;
  pt.rho = rho + 2 * p0 * N;
  pt.phi = nullptr;
  pt.mean = SC ? 0.0 : mean_part[kSumBlocks];
  pt.cosmo_grav = cosmo_grav;
  double acc[1] = {0.0};
  pt.racc = acc;
  pencil_stages<-1, LN, 0, kFromSource, kToLds, S, true, SC>(pt, pitch, tw, nullptr, G, 0, 0, 0);
  for (int e = threadIdx.x; e < TK * NH; e += kPencilThreads) {
    const int t = e / NH, k = e - t * NH;
    if (t < pt.cols) {
      const double2 zk = X[TI::at(k, t)], zn = X[TI::at((N - k) & (N - 1), t)];
      const double2 a = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
      const double2 b = make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
      double2 *ra = pt.base + (long long)(2 * t) * pitch + k;
      __builtin_nontemporal_store(a.x, &ra[0].x);
      __builtin_nontemporal_store(a.y, &ra[0].y);
      __builtin_nontemporal_store(b.x, &ra[pitch].x);
      __builtin_nontemporal_store(b.y, &ra[pitch].y);
    }
  }
  if constexpr (!SC) {  // the block's total of rho: a butterfly per wave, then the waves' totals
    __shared__ double wsum[kPencilThreads / 64];
    double v = acc[0];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < kPencilThreads / 64; i++) t += wsum[i];
      rho_sum[blockIdx.x] = t;
    }
  }
}

// k_xinv: the inverse (unnormalised, as a C2R transform) of the two rows'
// half spectra A and B as the one complex row A + i B; its real and
// imaginary parts are phi of rows 2t and 2t+1.
template <int LN>
__global__ void __launch_bounds__(kPencilThreads) __attribute__((amdgpu_waves_per_eu(kPencilWaves, kPencilWaves)))
k_xinv(const double2 *__restrict__ spec, double *__restrict__ phi, long long pitch, long long npairs,
       const double2 *__restrict__ tw, GParams G) {
  constexpr int N = 1 << LN, TK = 1 << (kPencilLog - LN);
  __shared__ double2 X[tile_elems<true>()];
  PencilTile pt;
  pt.X = X;
  const long long p0 = (long long)blockIdx.x * TK;
  pt.cols = (int)min((long long)TK, npairs - p0);
  pt.kx0 = 0;
  pt.outer = 0;
  pt.base = const_cast<double2 *>(spec) + 2 * p0 * pitch;  // read only: the last stage writes phi
  pt.rho = nullptr;
  pt.phi = phi + 2 * p0 * N;
  pt.mean = pt.cosmo_grav = 0.0;
  pencil_stages<1, LN, 0, kFromHalves, kToPhi, double, true>(pt, pitch, tw, nullptr, G, 0, 0, 0);
}

// ---- multigreen on the half spectrum (kx = 0..nx/2) after hipFFT's 3-D transform
// sin2 holds sin^2(kf2 kvalue) for x (nx), y (ny), z (nz)
__global__ void __launch_bounds__(256) k_green(double2 *__restrict__ spec, int nx, int ny, int nz,
                                               const double *__restrict__ sin2, GParams G) {
  const int nxh = nx / 2 + 1;
  const long long total = (long long)nxh * ny * nz;
  for (long long idx = blockIdx.x * 256LL + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int i = (int)(idx % nxh);
    const long long r = idx / nxh;
    const int j = (int)(r % ny), k = (int)(r / ny);
    spec[idx] = green_apply<true>(spec[idx], i, j, k, nx, ny, nz, sin2, G);
  }
}

// next-step limits of a cell from its stored q: fast or the reference's arithmetic
template <bool EXACT>
__device__ __forceinline__ int cell_limits(const KParams &P, double q[6], double &v, double &t) {
  double w[5];
  if (EXACT) {
    const int e = q2w(P, q, w);
    v = cfl_cell(P, w);
    t = w[4] / w[0] * P.temp0;
    return e;
  }
  double rinv;
  const int e = fast::q2w(P, q, w, rinv);
  v = fast::cfl_cell(P, w, rinv);
  t = w[4] * rinv * P.temp0;
  return e;
}

// ---- gravintegration + grav_force (gravitation.f90:172-254)
template <typename S, bool EXACT>
__global__ void __launch_bounds__(GT) k_gravint(KParams P, GParams Gp, GGeo G, S *__restrict__ q,
                                                const double *__restrict__ phi, double dt, Status *st,
                                                S *__restrict__ tout) {
  __shared__ double F[GRY * GRX];
  const int lin = xcd_tile(blockIdx.x, G.ntiles);
  if (lin < 0) return;  // padding block (whole block exits before any barrier)
  const int tile = lin % (G.tiles_x * G.tiles_y), chunk = lin / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * GX, y0 = (tile / G.tiles_x) * GY;
  const int z0 = chunk * G.zchunk, z1 = min(z0 + G.zchunk, G.nz);
  const int tid = threadIdx.x, cx = tid % GX, cy = tid / GX;
  const int gx = x0 + cx, gy = y0 + cy;
  const bool inside = gx < G.nx && gy < G.ny;
  int rx = 0, ry = 0;
  const bool ring = tid < GRING;
  if (tid < GX) { rx = tid; ry = -1; }
  else if (tid < 2 * GX) { rx = tid - GX; ry = GY; }
  else if (tid < 2 * GX + GY) { rx = -1; ry = tid - 2 * GX; }
  else if (tid < GRING) { rx = GX; ry = tid - 2 * GX - GY; }
  // periodic source cells (grav_force wraps in every direction); lanes of a
  // partial tile read clamped cells and update nothing
  const int sx = wrap(min(gx, G.nx), G.nx), sy = wrap(min(gy, G.ny), G.ny);
  const int rsx = wrap(min(x0 + rx, G.nx), G.nx), rsy = wrap(min(y0 + ry, G.ny), G.ny);
  const long long plane = (long long)G.nx * G.ny;
  const long long col = sx + (long long)G.nx * sy, rcol = rsx + (long long)G.nx * rsy;

  double pm = phi[col + plane * wrap(z0 - 1, G.nz)];
  double pc = phi[col + plane * z0];
  double vmin = 1e6;
  // qn: q of plane k, loaded one plane ahead (q(6) enters only the
  // dual-energy branch of the primitive below)
  S qn[6];
#pragma unroll
  for (int m = 0; m < 5; m++) qn[m] = q[m * G.n + col + plane * z0];
  qn[5] = P.dualenergy ? q[5 * G.n + col + plane * z0] : S(0);
  for (int k = z0; k < z1; k++) {
    double qc[6];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(qn[m]);
    if (k + 1 < z1) {
#pragma unroll
      for (int m = 0; m < 5; m++) qn[m] = q[m * G.n + col + plane * (k + 1)];
      if (P.dualenergy) qn[5] = q[5 * G.n + col + plane * (k + 1)];
    }
    const double pn = phi[col + plane * wrap(k + 1, G.nz)];
    F[(cx + 1) + GRX * (cy + 1)] = pc;
    if (ring) F[(rx + 1) + GRX * (ry + 1)] = phi[rcol + plane * k];
    __syncthreads();
    if (inside) {
      const int me = (cx + 1) + GRX * (cy + 1);
      double f[3];
      f[0] = G.nx == 1 ? 0.0 : -(F[me + 1] - F[me - 1]) * Gp.dx2inv;
      f[1] = G.ny == 1 ? 0.0 : -(F[me + GRX] - F[me - GRX]) * Gp.dy2inv;
      f[2] = G.nz == 1 ? 0.0 : -(pn - pm) * Gp.dz2inv;
      const long long c = col + plane * k;
      qc[4] = qc[4] + dt * (qc[1] * f[0] + qc[2] * f[1] + qc[3] * f[2]);
#pragma unroll
      for (int d = 0; d < 3; d++) qc[1 + d] = qc[1 + d] + dt * qc[0] * f[d];
#pragma unroll
      for (int m = 1; m < 5; m++) q[m * G.n + c] = S(qc[m]);
      // the next step's limits from the new state (as stored)
      double qs[6], v, t;
#pragma unroll
      for (int m = 0; m < 6; m++) qs[m] = double(S(qc[m]));
      const int e = cell_limits<EXACT>(P, qs, v, t);
      if (e) report_next(st, e, (unsigned long long)c);
      vmin = dmin(vmin, v);
      if (tout) tout[c] = S(t);
    }
    __syncthreads();  // F is rewritten for the next plane only after every lane has read it
    pm = pc;
    pc = pn;
  }
  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((tid & 63) == 0) atomic_min_pos(&st->vmin_bits, vmin);
}

GParams make_gparams(const evh_ctx *h) {
  const evh_config &c = h->cfg;
  GParams g;
  g.grav = c.grav;
  g.fb = c.fb;
  g.h2 = c.h2;
  g.n3 = double((long long)c.nx * c.ny * c.nz);
  g.n3inv = 1.0 / g.n3;
  g.dx2inv = c.dx2inv;
  g.dy2inv = c.dy2inv;
  g.dz2inv = c.dz2inv;
  g.supercomoving = c.supercomoving;
  return g;
}

GGeo make_ggeo(const evh_ctx *h) {
  GGeo g;
  g.nx = h->grid.nx;
  g.ny = h->grid.ny;
  g.nz = h->grid.nz;
  g.n = h->grid.n();
  g.tiles_x = (g.nx + GX - 1) / GX;
  g.tiles_y = (g.ny + GY - 1) / GY;
  g.zchunk = z_chunk((long long)g.tiles_x * g.tiles_y, g.nz, kZChunk);
  if (g.zchunk > g.nz) g.zchunk = g.nz;
  g.ntiles = g.tiles_x * g.tiles_y * ((g.nz + g.zchunk - 1) / g.zchunk);
  return g;
}

int grid_of(long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); }

// comp.f90:241-252
double kvalue(int i, int n, int nyq) { return i <= nyq ? double(i - 1) : -double(n - (i - 1)); }

hipError_t fft_status(hipfftResult r) { return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown; }

// With the pencil kernel, each row of the half spectrum (nx/2+1 values) is
// padded to a multiple of 16 values (256 B), so that the rows, and the
// pencil tiles of kx columns, start on cache-line boundaries: with the dense
// odd pitch (257 values at nx = 512) most row segments straddle two lines:
// the pencils read twice their bytes, and hipFFT's y pass slows as well.
constexpr int kSpecAlign = 16;

// lengths the pencil kernel takes: powers of two up to a block's tile
bool pencil_length(int n) { return n >= 2 && n <= kPencilElems && (n & (n - 1)) == 0; }

Pencils make_pencils(int n, int nxh, long long s_elem, long long s_outer) {
  Pencils p;
  p.ln = 0;
  while ((1 << p.ln) < n) p.ln++;
  const int tk = kPencilElems / n;  // kx columns per block
  p.nxh = nxh;
  p.ktiles = (nxh + tk - 1) / tk;
  p.s_elem = s_elem;
  p.s_outer = s_outer;
  return p;
}

// k_pencils instantiated for log2 n = LN .. kPencilLog; one block per tile
// of kx columns and outer index
template <int MODE, int LN = 1>
void launch_pencils(const Pencils &pc, int nouter, hipStream_t st, double2 *spec, const double2 *tw,
                    const double *sin2, const GParams &gp, int nx, int ny, int nz) {
  if constexpr (LN <= kPencilLog) {
    if (pc.ln == LN)
      hipLaunchKernelGGL((k_pencils<LN, MODE>), dim3(nouter * pc.ktiles), dim3(kPencilThreads), 0, st, spec, pc,
                         tw, sin2, gp, nx, ny, nz);
    else
      launch_pencils<MODE, LN + 1>(pc, nouter, st, spec, tw, sin2, gp, nx, ny, nz);
  }
}

// k_xfwd / k_xinv instantiated for log2 nx = LN .. kPencilLog; one block per
// tile of row pairs
template <typename S, int LN = 1>
void launch_xfwd(int ln, const S *rho, double2 *spec, long long pitch, long long npairs, hipStream_t st,
                 const double2 *tw, const GParams &gp, double cosmo_grav, const double *mean_part, double *rho_sum) {
  if constexpr (LN <= kPencilLog) {
    constexpr int TK = 1 << (kPencilLog - LN);
    const dim3 grid((unsigned)((npairs + TK - 1) / TK));
    if (ln == LN && gp.supercomoving)
      hipLaunchKernelGGL((k_xfwd<S, LN, true>), grid, dim3(kPencilThreads), 0, st, rho, spec, pitch, npairs, tw, gp,
                         cosmo_grav, mean_part, rho_sum);
    else if (ln == LN)
      hipLaunchKernelGGL((k_xfwd<S, LN, false>), grid, dim3(kPencilThreads), 0, st, rho, spec, pitch, npairs, tw, gp,
                         cosmo_grav, mean_part, rho_sum);
    else
      launch_xfwd<S, LN + 1>(ln, rho, spec, pitch, npairs, st, tw, gp, cosmo_grav, mean_part, rho_sum);
  }
}
// blocks of the x pass (the tiles of row pairs), one density total each
int xfwd_blocks(int nx, long long npairs) {
  int ln = 0;
  while ((1 << ln) < nx) ln++;
  const long long tk = 1LL << (kPencilLog - ln);
  return (int)((npairs + tk - 1) / tk);
}
template <int LN = 1>
void launch_xinv(int ln, const double2 *spec, double *phi, long long pitch, long long npairs, hipStream_t st,
                 const double2 *tw, const GParams &gp) {
  if constexpr (LN <= kPencilLog) {
    constexpr int TK = 1 << (kPencilLog - LN);
    if (ln == LN)
      hipLaunchKernelGGL((k_xinv<LN>), dim3((unsigned)((npairs + TK - 1) / TK)), dim3(kPencilThreads), 0, st, spec,
                         phi, pitch, npairs, tw, gp);
    else
      launch_xinv<LN + 1>(ln, spec, phi, pitch, npairs, st, tw, gp);
  }
}

// exp(-2 pi i t / n), t < n, in long double rounded once
void append_twiddles(std::vector<double2> &tw, int n) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; t++) {
    const long double a = two_pi * (long double)t / (long double)n;
    tw.push_back(make_double2((double)std::cos(a), -(double)std::sin(a)));
  }
}

template <typename S, bool EXACT>
void launch_gravint_t(evh_ctx *h, double dt) {
  const GGeo g = make_ggeo(h);
  const int blocks = xcd_blocks(g.ntiles);
  hipLaunchKernelGGL((k_gravint<S, EXACT>), dim3(blocks), dim3(GT), 0, h->stream, h->kp, make_gparams(h), g,
                     (S *)h->qbuf[h->cur], h->phi, dt, h->st, (S *)h->tbuf);
}

}  // namespace

hipError_t gravity_init(evh_ctx *h) {
  const evh_config &c = h->cfg;
  const long long n = h->grid.n();
  const bool pencils = pencil_length(c.nz);
  const bool xyz = pencils && pencil_length(c.ny) && pencil_length(c.nx);
  const int nxh = c.nx / 2 + 1;
  const int pitch = pencils ? (nxh + kSpecAlign - 1) / kSpecAlign * kSpecAlign : nxh;
  const long long plane = (long long)pitch * c.ny;
  const long long nspec = plane * c.nz;
  hipError_t e;
  if ((e = hipMalloc(&h->phi, sizeof(double) * n)) != hipSuccess) return e;
  if ((e = hipMalloc(&h->spec, sizeof(double2) * nspec)) != hipSuccess) return e;
  if ((e = hipMalloc(&h->sin2, sizeof(double) * (c.nx + c.ny + c.nz))) != hipSuccess) return e;
  if ((e = hipMalloc(&h->rho_part, sizeof(double) * (kSumBlocks + 1))) != hipSuccess) return e;
  h->dev_bytes += (long long)sizeof(double) * (n + 2 * nspec + c.nx + c.ny + c.nz + kSumBlocks + 1);
  if (xyz) {  // the x pass's block totals of rho
    h->rho_nsums = xfwd_blocks(c.nx, (long long)c.ny * c.nz / 2);
    if ((e = hipMalloc(&h->rho_sums, sizeof(double) * h->rho_nsums)) != hipSuccess) return e;
    h->dev_bytes += (long long)sizeof(double) * h->rho_nsums;
  }
  h->rho_sums_valid = false;
  // the sin^2 factors of multigreen, evaluated on the host like the reference
  std::vector<double> tab;
  const int ns[3] = {c.nx, c.ny, c.nz}, nyq[3] = {c.nyqx, c.nyqy, c.nyqz};
  const double kf2[3] = {c.kfx2, c.kfy2, c.kfz2};
  for (int d = 0; d < 3; d++)
    for (int i = 1; i <= ns[d]; i++) {
      const double s = std::sin(kf2[d] * kvalue(i, ns[d], nyq[d]));
      tab.push_back(s * s);
    }
  if ((e = hipMemcpy(h->sin2, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  hipfftHandle pf = nullptr, pi = nullptr;
  h->pencils = pencils;
  h->xyz_pencils = xyz;
  h->spec_plane = plane;
  h->spec_pitch = pitch;
  if (h->pencils) {
    // the x and y transforms: the kernels here when nx and ny are powers of
    // two too, else hipFFT's 2-D transforms, one per z plane (rows of nx
    // reals to rows of nx/2+1 values)
    if (!xyz) {
      int len[2] = {c.ny, c.nx}, rem[2] = {c.ny, c.nx}, cem[2] = {c.ny, pitch};
      const int rdist = c.ny * c.nx, cdist = (int)plane;
      if ((e = fft_status(hipfftPlanMany(&pf, 2, len, rem, 1, rdist, cem, 1, cdist, HIPFFT_D2Z, c.nz))) !=
          hipSuccess)
        return e;
      h->plan_fwd = pf;
      if ((e = fft_status(hipfftPlanMany(&pi, 2, len, cem, 1, cdist, rem, 1, rdist, HIPFFT_Z2D, c.nz))) !=
          hipSuccess)
        return e;
      h->plan_inv = pi;
    }
    // the twiddles of the z transforms, then of y and x
    std::vector<double2> tw;
    append_twiddles(tw, c.nz);
    if (xyz) {
      append_twiddles(tw, c.ny);
      append_twiddles(tw, c.nx);
    }
    if ((e = hipMalloc(&h->twiddle, sizeof(double2) * tw.size())) != hipSuccess) return e;
    h->dev_bytes += (long long)(sizeof(double2) * tw.size());
    if ((e = hipMemcpy(h->twiddle, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return e;
  } else if (c.nz > 1) {
    if ((e = fft_status(hipfftPlan3d(&pf, c.nz, c.ny, c.nx, HIPFFT_D2Z))) != hipSuccess) return e;
    h->plan_fwd = pf;
    if ((e = fft_status(hipfftPlan3d(&pi, c.nz, c.ny, c.nx, HIPFFT_Z2D))) != hipSuccess) return e;
    h->plan_inv = pi;
  } else {
    // 2-D and 1-D grids: the transform over the axes longer than one cell
    // (a length-1 axis of the reference's 3-D transform is the identity);
    // the half spectrum keeps its [nz][ny][nx/2+1] layout
    if (c.ny > 1) e = fft_status(hipfftPlan2d(&pf, c.ny, c.nx, HIPFFT_D2Z));
    else e = fft_status(hipfftPlan1d(&pf, c.nx, HIPFFT_D2Z, 1));
    if (e != hipSuccess) return e;
    h->plan_fwd = pf;
    if (c.ny > 1) e = fft_status(hipfftPlan2d(&pi, c.ny, c.nx, HIPFFT_Z2D));
    else e = fft_status(hipfftPlan1d(&pi, c.nx, HIPFFT_Z2D, 1));
    if (e != hipSuccess) return e;
    h->plan_inv = pi;
  }
  if (pf && (e = fft_status(hipfftSetStream(pf, h->stream))) != hipSuccess) return e;
  return pi ? fft_status(hipfftSetStream(pi, h->stream)) : hipSuccess;
}

void gravity_release(evh_ctx *h) {
  if (h->plan_fwd) hipfftDestroy((hipfftHandle)h->plan_fwd);
  if (h->plan_inv) hipfftDestroy((hipfftHandle)h->plan_inv);
  if (h->phi) hipFree(h->phi);
  if (h->spec) hipFree(h->spec);
  if (h->sin2) hipFree(h->sin2);
  if (h->rho_part) hipFree(h->rho_part);
  if (h->rho_sums) hipFree(h->rho_sums);
  h->rho_sums = nullptr;
  h->rho_nsums = 0;
  h->rho_sums_valid = false;
  if (h->twiddle) hipFree(h->twiddle);
  h->twiddle = nullptr;
  h->plan_fwd = h->plan_inv = nullptr;
  h->phi = nullptr;
  h->spec = nullptr;
  h->sin2 = h->rho_part = nullptr;
}

hipError_t launch_gravpotential(evh_ctx *h, double cosmo) {
  const long long n = h->grid.n();
  const GParams gp = make_gparams(h);
  const bool f32 = h->cfg.precision == EVH_FP32_MIXED;
  // The mean density only sets the (0,0,0) mode of the source, which the
  // Green's function zeroes; it keeps the transform's input small.  The
  // x pass of this library's transforms sums the densities it reads, and a
  // later call takes the mean of that earlier state (the same to rounding
  // while the flow conserves mass); the first call after an upload sums the
  // density first (gravitation.f90:39-42).
  const bool reuse = h->xyz_pencils && h->rho_sums_valid;
  if (!gp.supercomoving && reuse) {
    hipLaunchKernelGGL(k_rho_mean_from, dim3(1), dim3(kSumBlocks), 0, h->stream, h->rho_sums, h->rho_nsums, gp.n3,
                       h->rho_part);
  } else if (!gp.supercomoving) {
    if (f32)
      hipLaunchKernelGGL(k_rho_partials<float>, dim3(kSumBlocks), dim3(256), 0, h->stream,
                         (const float *)h->qbuf[h->cur], n, h->rho_part);
    else
      hipLaunchKernelGGL(k_rho_partials<double>, dim3(kSumBlocks), dim3(256), 0, h->stream,
                         (const double *)h->qbuf[h->cur], n, h->rho_part);
    hipLaunchKernelGGL(k_rho_mean, dim3(1), dim3(kSumBlocks), 0, h->stream, h->rho_part, gp.n3);
  }
  const double cosmo_grav = cosmo * gp.grav;
  const int nx = h->grid.nx, ny = h->grid.ny, nz = h->grid.nz, nxh = nx / 2 + 1;
  if (h->xyz_pencils) {
    // source and x transform, y, z with the Green's function, y, x to phi
    double2 *spec = (double2 *)h->spec;
    const double2 *twz = (const double2 *)h->twiddle, *twy = twz + nz, *twx = twy + ny;
    const Pencils pz = make_pencils(nz, nxh, h->spec_plane, h->spec_pitch);
    const Pencils py = make_pencils(ny, nxh, h->spec_pitch, h->spec_plane);
    const long long npairs = (long long)ny * nz / 2;
    int lnx = 0;
    while ((1 << lnx) < nx) lnx++;
    if (f32)
      launch_xfwd<float>(lnx, (const float *)h->qbuf[h->cur], spec, h->spec_pitch, npairs, h->stream, twx, gp,
                         cosmo_grav, h->rho_part, h->rho_sums);
    else
      launch_xfwd<double>(lnx, (const double *)h->qbuf[h->cur], spec, h->spec_pitch, npairs, h->stream, twx, gp,
                          cosmo_grav, h->rho_part, h->rho_sums);
    h->rho_sums_valid = !gp.supercomoving;
    launch_pencils<kPencilFwdY>(py, nz, h->stream, spec, twy, h->sin2, gp, nx, ny, nz);
    launch_pencils<kPencilGreenZ>(pz, ny, h->stream, spec, twz, h->sin2, gp, nx, ny, nz);
    launch_pencils<kPencilInvY>(py, nz, h->stream, spec, twy, h->sin2, gp, nx, ny, nz);
    launch_xinv(lnx, spec, h->phi, h->spec_pitch, npairs, h->stream, twx, gp);
    return hipGetLastError();
  }
  if (f32)
    hipLaunchKernelGGL(k_grav_source<float>, dim3(grid_of(n)), dim3(256), 0, h->stream,
                       (const float *)h->qbuf[h->cur], n, h->phi, gp, cosmo_grav, h->rho_part);
  else
    hipLaunchKernelGGL(k_grav_source<double>, dim3(grid_of(n)), dim3(256), 0, h->stream,
                       (const double *)h->qbuf[h->cur], n, h->phi, gp, cosmo_grav, h->rho_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = fft_status(hipfftExecD2Z((hipfftHandle)h->plan_fwd, h->phi, (hipfftDoubleComplex *)h->spec))) !=
      hipSuccess)
    return e;
  if (h->pencils) {
    const Pencils pz = make_pencils(nz, nxh, h->spec_plane, h->spec_pitch);
    launch_pencils<kPencilGreenZ>(pz, ny, h->stream, (double2 *)h->spec, (const double2 *)h->twiddle, h->sin2, gp,
                                  nx, ny, nz);
  } else {
    const long long nspec = (long long)nxh * ny * nz;
    hipLaunchKernelGGL(k_green, dim3(grid_of(nspec)), dim3(256), 0, h->stream, (double2 *)h->spec, nx, ny, nz,
                       h->sin2, gp);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return fft_status(hipfftExecZ2D((hipfftHandle)h->plan_inv, (hipfftDoubleComplex *)h->spec, h->phi));
}

// the caller resets the CFL minimum before
hipError_t launch_gravintegration(evh_ctx *h, double dt) {
  const bool exact = h->mode != EVH_MODE_FUSED && h->cfg.precision == EVH_FP64;
  if (h->cfg.precision == EVH_FP32_MIXED) launch_gravint_t<float, false>(h, dt);
  else if (exact) launch_gravint_t<double, true>(h, dt);
  else launch_gravint_t<double, false>(h, dt);
  return hipGetLastError();
}

}  // namespace evh
