// staged.hip -- one kernel per reference routine, FP64, arrays in HBM.
//
// This is the reference-structured form of the step: primitive writes w,
// bound fills the ghosts, hydrostep reduces the CFL limit, compevo writes
// evo and s, and the sweep applies the x, then y, then z flux differences to
// q.  It serves 1-D and 2-D grids (whose compevo association differs,
// muscl.f90:60-88) and checks each routine on its own against the oracle.
// Built with -ffp-contract=off: results equal the reference's bit for bit
// except for the pow() feeding the passive entropy component.
#include "evh_internal.h"

namespace evh {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ long long widx(const Grid &G, int i, int j, int k) {  // reference indices, ghosts -2..n+1 (0-based)
  return (long long)(i + 2) + (long long)(G.nx + 4) * ((long long)(j + 2) + (long long)(G.ny + 4) * (k + 2));
}
__device__ __forceinline__ long long eidx(const Grid &G, int i, int j, int k) {  // 0-based cell i in -1..n
  return (long long)(i + 1) + (long long)(G.nx + 2) * ((long long)(j + 1) + (long long)(G.ny + 2) * (k + 1));
}

// ------------------------------------------------------------ primitive (muscl.f90:26-43)
__global__ void __launch_bounds__(kBlock) k_primitive(KParams P, Grid G, double *__restrict__ q,
                                                      double *__restrict__ w, Status *st) {
  const long long n = G.n();
  for (long long c = blockIdx.x * (long long)kBlock + threadIdx.x; c < n; c += (long long)gridDim.x * kBlock) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((long long)G.nx * G.ny));
    double qc[6], wc[5];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = q[m * n + c];
    const int e = q2w(P, qc, wc);
    if (e) report(st, e, (unsigned long long)c);
    const long long wi = widx(G, i, j, k);
    const long long nw = G.nw();
#pragma unroll
    for (int m = 0; m < 5; m++) w[m * nw + wi] = wc[m];
    if (P.dualenergy) {
      q[4 * n + c] = qc[4];
      q[5 * n + c] = qc[5];
    }
  }
}

// ------------------------------------------------------------ bound (boundary.f90:26-113)
// AXIS 0: x ghosts over all (j,k) including ghosts; 1: y; 2: z.
template <int AXIS>
__global__ void __launch_bounds__(kBlock) k_bound(KParams P, Grid G, double *__restrict__ w) {
  const int ea = AXIS == 0 ? G.ny + 4 : G.nx + 4;  // first free extent
  const int eb = AXIS == 2 ? G.ny + 4 : G.nz + 4;  // second free extent
  const long long total = (long long)ea * eb * 4;
  const long long nw = G.nw();
  const int n = AXIS == 0 ? G.nx : (AXIS == 1 ? G.ny : G.nz);
  for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < total; t += (long long)gridDim.x * kBlock) {
    const int slot = (int)(t % 4);
    const int a = (int)((t / 4) % ea) - 2;
    const int b = (int)(t / (4LL * ea)) - 2;
    // ghost positions: -2, -1, n, n+1 (0-based)
    const int g = slot < 2 ? slot - 2 : n + (slot - 2);
    const bool upper = slot >= 2;
    int src;
    bool negate = false;
    if (P.boundtype == 2) {
      src = g < 0 ? g + n : g - n;
    } else {
      src = g < 0 ? -1 - g : 2 * n - 1 - g;
      negate = (AXIS == 0 && upper && P.boundtype == 3);
    }
    long long dst_i, src_i;
    if (AXIS == 0) {
      dst_i = widx(G, g, a, b);
      src_i = widx(G, src, a, b);
    } else if (AXIS == 1) {
      dst_i = widx(G, a, g, b);
      src_i = widx(G, a, src, b);
    } else {
      dst_i = widx(G, a, b, g);
      src_i = widx(G, a, b, src);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const double v = w[m * nw + src_i];
      w[m * nw + dst_i] = (negate && m >= 1 && m <= 3) ? -v : v;
    }
  }
}

// ------------------------------------------------------------ hydrostep (timestep.f90:26-57)
__global__ void __launch_bounds__(kBlock) k_hydrostep(KParams P, Grid G, const double *__restrict__ w, Status *st) {
  const long long n = G.n();
  const long long nw = G.nw();
  double v = 1e6;
  for (long long c = blockIdx.x * (long long)kBlock + threadIdx.x; c < n; c += (long long)gridDim.x * kBlock) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((long long)G.nx * G.ny));
    const long long wi = widx(G, i, j, k);
    double wc[5];
#pragma unroll
    for (int m = 0; m < 5; m++) wc[m] = w[m * nw + wi];
    v = dmin(v, cfl_cell(P, wc));
  }
  // wavefront min, then one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) atomic_min_pos(&st->vmin_bits, v);
}

// ------------------------------------------------------------ compevo (muscl.f90:45-113)
template <int LIMITER, int DIM>
__global__ void __launch_bounds__(kBlock) k_compevo(KParams P, Grid G, const double *__restrict__ w,
                                                    double *__restrict__ evo, double *__restrict__ s, double dt) {
  const int ex = G.nx + 2;
  const int ey = DIM >= 2 ? G.ny + 2 : 1;
  const int ez = DIM == 3 ? G.nz + 2 : 1;
  const long long total = (long long)ex * ey * ez;
  const long long nw = G.nw(), ne = G.ne();
  const double dt2 = -0.5 * dt;
  for (long long t = blockIdx.x * (long long)kBlock + threadIdx.x; t < total; t += (long long)gridDim.x * kBlock) {
    const int i = (int)(t % ex) - 1;
    const int j = DIM >= 2 ? (int)((t / ex) % ey) - 1 : 0;
    const int k = DIM == 3 ? (int)(t / ((long long)ex * ey)) - 1 : 0;
    const long long c = widx(G, i, j, k);
    const long long sx = 1, sy = G.nx + 4, sz = (long long)(G.nx + 4) * (G.ny + 4);
    double wc[5], sl[3][5], a[3][5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      wc[m] = w[m * nw + c];
      sl[0][m] = slope3<LIMITER>(w[m * nw + c - sx], wc[m], w[m * nw + c + sx]);
      if (DIM >= 2) sl[1][m] = slope3<LIMITER>(w[m * nw + c - sy], wc[m], w[m * nw + c + sy]);
      if (DIM == 3) sl[2][m] = slope3<LIMITER>(w[m * nw + c - sz], wc[m], w[m * nw + c + sz]);
    }
    predictor_dir<0>(P, wc, sl[0], a[0]);
    if (DIM >= 2) predictor_dir<1>(P, wc, sl[1], a[1]);
    if (DIM == 3) predictor_dir<2>(P, wc, sl[2], a[2]);
    const long long e = eidx(G, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) {
      double ev;
      if (DIM == 1) ev = dt2 * a[0][m] * P.dxinv;                                   // muscl.f90:68
      else if (DIM == 2) ev = dt2 * (a[0][m] * P.dxinv + a[1][m] * P.dyinv);        // muscl.f90:86
      else ev = dt2 * (a[0][m] * P.dxinv + a[1][m] * P.dyinv + a[2][m] * P.dzinv);  // muscl.f90:106-107
      evo[m * ne + e] = ev;
#pragma unroll
      for (int d = 0; d < DIM; d++) s[(m + 5 * d) * ne + e] = sl[d][m];
    }
  }
}

// ------------------------------------------------------------ sweep (muscl.f90:171-274)
// One direction over all interior cells: both faces of the cell are solved
// (each face twice across the grid, with identical operands), then
// q += delta*(fl - fr) as integration (muscl.f90:287).
template <int SOLVER, int DIR>
__global__ void __launch_bounds__(kBlock) k_sweep(KParams P, Grid G, double *__restrict__ q,
                                                  const double *__restrict__ w, const double *__restrict__ evo,
                                                  const double *__restrict__ s, double delta, Status *st) {
  const long long n = G.n();
  const long long nw = G.nw(), ne = G.ne();
  const long long wst = DIR == 1 ? 1 : (DIR == 2 ? (G.nx + 4) : (long long)(G.nx + 4) * (G.ny + 4));
  const long long est = DIR == 1 ? 1 : (DIR == 2 ? (G.nx + 2) : (long long)(G.nx + 2) * (G.ny + 2));
  for (long long c = blockIdx.x * (long long)kBlock + threadIdx.x; c < n; c += (long long)gridDim.x * kBlock) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((long long)G.nx * G.ny));
    const long long wc = widx(G, i, j, k), ec = eidx(G, i, j, k);
    double f[2][6];
    int err = kOk;
#pragma unroll
    for (int side = 0; side < 2; side++) {  // face (c-1, c) then (c, c+1)
      const long long w0 = wc + (side - 1) * wst, e0 = ec + (side - 1) * est;
      double wl[5], wr[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        wl[m] = w[m * nw + w0];
        wr[m] = w[m * nw + w0 + wst];
      }
      if (P.secondorder) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
          wl[m] = wl[m] + 0.5 * s[(m + 5 * (DIR - 1)) * ne + e0] + evo[m * ne + e0];
          wr[m] = wr[m] - 0.5 * s[(m + 5 * (DIR - 1)) * ne + e0 + est] + evo[m * ne + e0 + est];
        }
      }
      const int e = riemann<SOLVER, DIR>(P, wl, wr, f[side]);
      if (e) err = e;
    }
    if (err) report(st, err, (unsigned long long)c);
#pragma unroll
    for (int m = 0; m < 6; m++) q[m * n + c] = q[m * n + c] + delta * (f[0][m] - f[1][m]);
  }
}

int grid_for(long long work) {
  long long b = (work + kBlock - 1) / kBlock;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

hipError_t launch_primitive(evh_ctx *h) {
  hipLaunchKernelGGL(k_primitive, dim3(grid_for(h->grid.n())), dim3(kBlock), 0, h->stream, h->kp, h->grid,
                     (double *)h->qbuf[h->cur], h->w, h->st);
  return hipGetLastError();
}

hipError_t launch_bound(evh_ctx *h) {
  const Grid &G = h->grid;
  if (G.nx != 1)
    hipLaunchKernelGGL(k_bound<0>, dim3(grid_for(4LL * (G.ny + 4) * (G.nz + 4))), dim3(kBlock), 0, h->stream, h->kp,
                       G, h->w);
  if (G.ny != 1)
    hipLaunchKernelGGL(k_bound<1>, dim3(grid_for(4LL * (G.nx + 4) * (G.nz + 4))), dim3(kBlock), 0, h->stream, h->kp,
                       G, h->w);
  if (G.nz != 1)
    hipLaunchKernelGGL(k_bound<2>, dim3(grid_for(4LL * (G.nx + 4) * (G.ny + 4))), dim3(kBlock), 0, h->stream, h->kp,
                       G, h->w);
  return hipGetLastError();
}

hipError_t launch_hydrostep(evh_ctx *h) {
  hipLaunchKernelGGL(k_hydrostep, dim3(grid_for(h->grid.n())), dim3(kBlock), 0, h->stream, h->kp, h->grid, h->w,
                     h->st);
  return hipGetLastError();
}

template <int LIMITER>
static void compevo_dim(evh_ctx *h, double dt) {
  const Grid &G = h->grid;
  const long long work = (long long)(G.nx + 2) * (h->kp.onedimension ? 1 : G.ny + 2) *
                         (h->kp.twodimension ? 1 : G.nz + 2);
  const dim3 gr(grid_for(work)), bl(kBlock);
  if (h->kp.onedimension)
    hipLaunchKernelGGL((k_compevo<LIMITER, 1>), gr, bl, 0, h->stream, h->kp, G, h->w, h->evo, h->s, dt);
  else if (h->kp.twodimension)
    hipLaunchKernelGGL((k_compevo<LIMITER, 2>), gr, bl, 0, h->stream, h->kp, G, h->w, h->evo, h->s, dt);
  else
    hipLaunchKernelGGL((k_compevo<LIMITER, 3>), gr, bl, 0, h->stream, h->kp, G, h->w, h->evo, h->s, dt);
}

hipError_t launch_compevo(evh_ctx *h, double dt) {
  switch (h->kp.limiter) {
    case 1: compevo_dim<1>(h, dt); break;
    case 2: compevo_dim<2>(h, dt); break;
    case 3: compevo_dim<3>(h, dt); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int SOLVER>
static void sweep_solver(evh_ctx *h, double dt) {
  const Grid &G = h->grid;
  const dim3 gr(grid_for(G.n())), bl(kBlock);
  double *q = (double *)h->qbuf[h->cur];
  if (G.nx != 1)
    hipLaunchKernelGGL((k_sweep<SOLVER, 1>), gr, bl, 0, h->stream, h->kp, G, q, h->w, h->evo, h->s, dt * h->kp.dxinv,
                       h->st);
  if (G.ny != 1)
    hipLaunchKernelGGL((k_sweep<SOLVER, 2>), gr, bl, 0, h->stream, h->kp, G, q, h->w, h->evo, h->s, dt * h->kp.dyinv,
                       h->st);
  if (G.nz != 1)
    hipLaunchKernelGGL((k_sweep<SOLVER, 3>), gr, bl, 0, h->stream, h->kp, G, q, h->w, h->evo, h->s, dt * h->kp.dzinv,
                       h->st);
}

hipError_t launch_sweep(evh_ctx *h, double dt) {
  if (h->kp.solver == 1) sweep_solver<1>(h, dt);
  else if (h->kp.solver == 2) sweep_solver<2>(h, dt);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace evh
