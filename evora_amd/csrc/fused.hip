// fused.hip -- the whole 3-D hydro step of one grid in one HBM pass.
//
// k_fused runs the tile march of fused_core.h over one grid: one block per
// (16 x TY tile, z chunk), with the tiles of each XCD contiguous; on 2-D
// grids (nz = 1), k_fused2d marches row strips along y (march2d.h).  k_cfl
// gives the first step's CFL limit.  The file is compiled three times:
// EVH_FUSED_PREC=64 (FP64, fast arithmetic), EVH_FUSED_PREC=32 (FP32-mixed)
// and EVH_FUSED_EXACT=1 (FP64 with the reference's arithmetic, compiled
// without FMA contraction).
#include "fused_core.h"
#include "march2d.h"

#ifndef EVH_FUSED_PREC
#define EVH_FUSED_PREC 64
#endif
#ifndef EVH_FUSED_EXACT
#define EVH_FUSED_EXACT 0
#endif
#if EVH_FUSED_EXACT && EVH_FUSED_PREC != 64
#error "the exact form is FP64"
#endif
constexpr bool kExact = EVH_FUSED_EXACT != 0;
// the exact build's kernel has a name of its own, so that profiles keep the
// two FP64 forms apart
#if EVH_FUSED_EXACT
#define K_FUSED k_fused_exact
#define K_FUSED2D k_fused2d_exact
#else
#define K_FUSED k_fused
#define K_FUSED2D k_fused2d
#endif

namespace evh {
namespace {

using core::Geo;

// WPE: waves per SIMD the register allocation must leave room for
// (512 / WPE VGPRs per lane); the tuning knob next to the tile height TY.
// SLAB: a z-slab whose ghost planes may come from halo buffers.
template <typename T, typename S, int SOLVER, int LIMITER, int TY, int WPE, bool SLAB>
__global__ void __launch_bounds__(core::TileShape<TY>::NT) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    K_FUSED(KParams P, Geo G, const S *__restrict__ qin, S *__restrict__ qout, double dt, Status *st,
            S *__restrict__ tout) {
  __shared__ core::Smem<T, TY> sm;
  const int lin = xcd_tile(blockIdx.x, G.ntiles);
  if (lin < 0) return;  // padding block (whole block exits before any barrier)
  core::march<T, S, SOLVER, LIMITER, TY, kExact, SLAB>(P, G, qin, qout, 0, 0, dt, st, &st->vmin_bits, lin, sm,
                                                        tout);
}

// 2-D grids (nz = 1): row strips marching along y (march2d.h)
template <typename T, typename S, int SOLVER, int LIMITER, int WPE>
__global__ void __launch_bounds__(core2d::NL) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    K_FUSED2D(KParams P, core2d::Geo2 G, const S *__restrict__ qin, S *__restrict__ qout, double dt, Status *st,
              S *__restrict__ tout) {
  __shared__ core2d::Smem2<T> sm;
  const int lin = xcd_tile(blockIdx.x, G.ntiles);
  if (lin < 0) return;  // padding block (whole block exits before any barrier)
  core2d::march2d<T, S, SOLVER, LIMITER, kExact>(P, G, qin, qout, dt, st, &st->vmin_bits, lin, sm, tout);
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

Geo make_geo(const evh_ctx *h, int ty) {
  Geo g;
  g.nx = h->grid.nx;
  g.ny = h->grid.ny;
  g.nz = h->grid.nz;
  g.n = h->grid.n();
  g.tiles_x = (g.nx + core::TX - 1) / core::TX;
  g.tiles_y = (g.ny + ty - 1) / ty;
  g.zchunk = h->zchunk > 0 ? h->zchunk : z_chunk((long long)g.tiles_x * g.tiles_y, g.nz, 64);
  if (g.zchunk > g.nz) g.zchunk = g.nz;
  g.nchunks = (g.nz + g.zchunk - 1) / g.zchunk;
  g.ntiles = g.tiles_x * g.tiles_y * g.nchunks;
  g.zlo = h->zslab[0] ? h->zhalo[0] : nullptr;
  g.zhi = h->zslab[1] ? h->zhalo[1] : nullptr;
  return g;
}

template <typename T, typename S, int SOLVER, int LIMITER, int TY, int WPE, bool SLAB = false>
void launch_ty(evh_ctx *h, double dt) {
  const Geo g = make_geo(h, TY);
  const int blocks = xcd_blocks(g.ntiles);
  hipLaunchKernelGGL((K_FUSED<T, S, SOLVER, LIMITER, TY, WPE, SLAB>), dim3(blocks), dim3(core::TileShape<TY>::NT), 0,
                     h->stream, h->kp, g, (const S *)h->qbuf[h->cur], (S *)h->qbuf[1 - h->cur], dt, h->st,
                     (S *)h->tbuf);
}

// the 2-D form: strips of 252 columns, chunks of up to 32 rows (the z chunk
// of the tuning, when set), WPE waves per SIMD
template <typename T, typename S, int SOLVER, int LIMITER, int WPE>
void launch_2d(evh_ctx *h, double dt) {
  core2d::Geo2 g;
  g.nx = h->grid.nx;
  g.ny = h->grid.ny;
  g.n = h->grid.n();
  g.strips = (g.nx + core2d::NC - 1) / core2d::NC;
  // 32 rows measured fastest at 8192^2 (FP64 2.10 ms against 2.14 with 64 and
  // 2.27 with 128; FP32-mixed 0.94 with 32 or 64)
  g.ychunk = h->zchunk > 0 ? h->zchunk : y_chunk(g.strips, g.ny, 32);
  if (g.ychunk > g.ny) g.ychunk = g.ny;
  g.nchunks = (g.ny + g.ychunk - 1) / g.ychunk;
  g.ntiles = g.strips * g.nchunks;
  hipLaunchKernelGGL((K_FUSED2D<T, S, SOLVER, LIMITER, WPE>), dim3(xcd_blocks(g.ntiles)), dim3(core2d::NL), 0,
                     h->stream, h->kp, g, (const S *)h->qbuf[h->cur], (S *)h->qbuf[1 - h->cur], dt, h->st,
                     (S *)h->tbuf);
}

// tuning variants: tile height 8 or 12, waves per SIMD (FP64: 2 or 3, FP32: 3
// or 4); 2-D grids take the row-strip march in its one shape
template <typename T, typename S, int SOLVER, int LIMITER>
void launch_fused_t(evh_ctx *h, double dt) {
  constexpr int LO = sizeof(T) == 8 ? 2 : 3, HI = LO + 1;
  if (h->grid.nz == 1) {
    // 2-D: FP64 at 3 waves per SIMD (its row carry in LDS, 16 B/lane
    // spilled: 3% faster than 2 waves at 8192^2), FP32 at 4 without spills
    // (5 spill 32 B/lane and measured 4% slower)
    launch_2d<T, S, SOLVER, LIMITER, sizeof(T) == 8 ? 3 : 4>(h, dt);
    return;
  }
  if (h->zslab[0] || h->zslab[1]) {
    // a z-slab with halo buffers: the default tile shape (the z chunk is
    // still the tuning's)
    launch_ty<T, S, SOLVER, LIMITER, 12, HI, true>(h, dt);
    return;
  }
  const bool hi = h->wpe_hi;
  if (h->tile_ty == 8) {
    if (hi) launch_ty<T, S, SOLVER, LIMITER, 8, HI>(h, dt);
    else launch_ty<T, S, SOLVER, LIMITER, 8, LO>(h, dt);
  } else {
    if (hi) launch_ty<T, S, SOLVER, LIMITER, 12, HI>(h, dt);
    else launch_ty<T, S, SOLVER, LIMITER, 12, LO>(h, dt);
  }
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

#if EVH_FUSED_PREC == 64 && !EVH_FUSED_EXACT
// 3-D grids, and 2-D grids (nz = 1) in the march's 2-D form
bool fused_supported(const evh_config &c) { return c.nx >= 2 && c.ny >= 2 && c.nz >= 1; }

hipError_t launch_cfl(evh_ctx *h) {
  const Geo g = make_geo(h, 12);
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
#endif

#if EVH_FUSED_EXACT
hipError_t launch_fused_exact(evh_ctx *h, double dt) { return launch_fused_prec<double, double>(h, dt); }
#elif EVH_FUSED_PREC == 32
hipError_t launch_fused_f32(evh_ctx *h, double dt) { return launch_fused_prec<float, float>(h, dt); }
#else
hipError_t launch_fused_f64(evh_ctx *h, double dt) { return launch_fused_prec<double, double>(h, dt); }

hipError_t launch_fused_step(evh_ctx *h, double dt) {
  if (h->cfg.precision == EVH_FP32_MIXED) return launch_fused_f32(h, dt);
  if (h->mode == EVH_MODE_FUSED_EXACT) return launch_fused_exact(h, dt);
  return launch_fused_f64(h, dt);
}
#endif

}  // namespace evh
