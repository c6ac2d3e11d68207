// gravity.hip -- self-gravity on the device: the Poisson solve of
// gravpotential (gravitation.f90:26-88) and gravintegration with grav_force
// (gravitation.f90:172-254), one rank.
//
// Potential.  The source grav (rho - mean rho) (or, supercomoving,
// cosmo grav (rho / fb - 1)) goes through a real-to-complex FFT (hipFFT,
// FP64 in both precision forms), is multiplied by the Green's function of
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
  double grav, fb, h2, n3;
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
  double s = 0.0, c = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double y = double(q[i]) - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  red[threadIdx.x] = s;
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

// ---- multigreen (gravitation.f90:55-89) on the half spectrum (kx = 0..nx/2)
// sin2 holds sin^2(kf2 kvalue) for x (nx), y (ny), z (nz)
__global__ void __launch_bounds__(256) k_green(double2 *__restrict__ spec, int nx, int ny, int nz,
                                               const double *__restrict__ sin2, GParams G) {
  const int nxh = nx / 2 + 1;
  const long long total = (long long)nxh * ny * nz;
  const double *sx = sin2, *sy = sin2 + nx, *sz = sin2 + nx + ny;
  const double den = G.n3 * G.n3 + 0.0 * 0.0;  // phi / dble(n3) as the complex quotient
  for (long long idx = blockIdx.x * 256LL + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int i = (int)(idx % nxh);
    const long long r = idx / nxh;
    const int j = (int)(r % ny), k = (int)(r / ny);
    double green;
    if (i == 0 && j == 0 && k == 0) {
      green = 0.0;
    } else {
      const int im = i == 0 ? 0 : nx - i, jm = j == 0 ? 0 : ny - j, km = k == 0 ? 0 : nz - k;
      const double g1 = -G.h2 / (sx[i] + sy[j] + sz[k]);
      const double g2 = -G.h2 / (sx[im] + sy[jm] + sz[km]);
      green = g1 == g2 ? g1 : 0.5 * (g1 + g2);
    }
    const double2 v = spec[idx];
    double2 o;
    o.x = (v.x * G.n3 + v.y * 0.0) / den * green;
    o.y = (v.y * G.n3 - v.x * 0.0) / den * green;
    spec[idx] = o;
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
  for (int k = z0; k < z1; k++) {
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
      // q(6) enters only the dual-energy branch of the primitive below
      double qc[6];
#pragma unroll
      for (int m = 0; m < 5; m++) qc[m] = double(q[m * G.n + c]);
      qc[5] = P.dualenergy ? double(q[5 * G.n + c]) : 0.0;
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
  g.zchunk = g.nz < kZChunk ? g.nz : kZChunk;
  g.tiles_x = (g.nx + GX - 1) / GX;
  g.tiles_y = (g.ny + GY - 1) / GY;
  g.ntiles = g.tiles_x * g.tiles_y * ((g.nz + g.zchunk - 1) / g.zchunk);
  return g;
}

int grid_of(long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); }

// comp.f90:241-252
double kvalue(int i, int n, int nyq) { return i <= nyq ? double(i - 1) : -double(n - (i - 1)); }

hipError_t fft_status(hipfftResult r) { return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown; }

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
  const long long nspec = (long long)(c.nx / 2 + 1) * c.ny * c.nz;
  hipError_t e;
  if ((e = hipMalloc(&h->phi, sizeof(double) * n)) != hipSuccess) return e;
  if ((e = hipMalloc(&h->spec, sizeof(double2) * nspec)) != hipSuccess) return e;
  if ((e = hipMalloc(&h->sin2, sizeof(double) * (c.nx + c.ny + c.nz))) != hipSuccess) return e;
  if ((e = hipMalloc(&h->rho_part, sizeof(double) * (kSumBlocks + 1))) != hipSuccess) return e;
  h->dev_bytes += (long long)sizeof(double) * (n + 2 * nspec + c.nx + c.ny + c.nz + kSumBlocks + 1);
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
  if ((e = fft_status(hipfftPlan3d(&pf, c.nz, c.ny, c.nx, HIPFFT_D2Z))) != hipSuccess) return e;
  h->plan_fwd = pf;
  if ((e = fft_status(hipfftPlan3d(&pi, c.nz, c.ny, c.nx, HIPFFT_Z2D))) != hipSuccess) return e;
  h->plan_inv = pi;
  if ((e = fft_status(hipfftSetStream(pf, h->stream))) != hipSuccess) return e;
  return fft_status(hipfftSetStream(pi, h->stream));
}

void gravity_release(evh_ctx *h) {
  if (h->plan_fwd) hipfftDestroy((hipfftHandle)h->plan_fwd);
  if (h->plan_inv) hipfftDestroy((hipfftHandle)h->plan_inv);
  if (h->phi) hipFree(h->phi);
  if (h->spec) hipFree(h->spec);
  if (h->sin2) hipFree(h->sin2);
  if (h->rho_part) hipFree(h->rho_part);
  h->plan_fwd = h->plan_inv = nullptr;
  h->phi = nullptr;
  h->spec = nullptr;
  h->sin2 = h->rho_part = nullptr;
}

hipError_t launch_gravpotential(evh_ctx *h, double cosmo) {
  const long long n = h->grid.n();
  const GParams gp = make_gparams(h);
  const bool f32 = h->cfg.precision == EVH_FP32_MIXED;
  if (!gp.supercomoving) {
    if (f32)
      hipLaunchKernelGGL(k_rho_partials<float>, dim3(kSumBlocks), dim3(256), 0, h->stream,
                         (const float *)h->qbuf[h->cur], n, h->rho_part);
    else
      hipLaunchKernelGGL(k_rho_partials<double>, dim3(kSumBlocks), dim3(256), 0, h->stream,
                         (const double *)h->qbuf[h->cur], n, h->rho_part);
    hipLaunchKernelGGL(k_rho_mean, dim3(1), dim3(kSumBlocks), 0, h->stream, h->rho_part, gp.n3);
  }
  const double cosmo_grav = cosmo * gp.grav;
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
  const long long nspec = (long long)(h->grid.nx / 2 + 1) * h->grid.ny * h->grid.nz;
  hipLaunchKernelGGL(k_green, dim3(grid_of(nspec)), dim3(256), 0, h->stream, (double2 *)h->spec, h->grid.nx,
                     h->grid.ny, h->grid.nz, h->sin2, gp);
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
