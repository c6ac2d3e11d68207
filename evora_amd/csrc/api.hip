// api.hip -- the C-ABI of libevora_hip (include/evora_hip.h).
//
// Host side of the drop-in: owns the device state of one grid, maps the
// reference's step-level calls (primitive, bound, hydrostep, compevo, sweep)
// onto either the staged kernels or the fused one-pass kernel, and converts
// between the host AoS layout q(6,nx,ny,nz) and the device SoA layout.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "evh_internal.h"

using namespace evh;

namespace {

constexpr double kVInit = 1e6;  // timestep.f90:35

const char *err_text(int code) {
  switch (code) {
    case EVH_OK: return "ok";
    case EVH_ERR_NEG_PRESSURE: return "ERROR negative pressure occured while computing primitive variables!";
    case EVH_ERR_RIEMANN_PRESSURE: return "ERROR: riemann solver unphysical pressure";
    case EVH_ERR_HLL: return "ERROR: hll solver";
    case EVH_ERR_HLLC: return "ERROR: hllc solver";
    case EVH_ERR_DT: return "ERROR: hydrostep is negative";
    case EVH_ERR_SOLVER: return "ERROR choose proper solver!";
    case EVH_ERR_LIMITER: return "ERROR: wrong slope limiter!";
    case EVH_ERR_CONFIG: return "ERROR: unsupported configuration";
    case EVH_ERR_HIP: return "ERROR: HIP runtime failure";
    case EVH_ERR_STATE: return "ERROR: call out of order";
    default: return "ERROR: unknown";
  }
}

int fail(evh_ctx *h, int code, const char *detail = nullptr, long long cell = -1) {
  if (!h) return code;
  h->err_code = code;
  h->err_msg = err_text(code);
  if (detail) {
    h->err_msg += " (";
    h->err_msg += detail;
    h->err_msg += ")";
  }
  if (cell >= 0) {
    const long long nx = h->grid.nx, ny = h->grid.ny;
    h->err_cell[0] = (int)(cell % nx) + 1;
    h->err_cell[1] = (int)((cell / nx) % ny) + 1;
    h->err_cell[2] = (int)(cell / (nx * ny)) + 1;
  } else {
    h->err_cell[0] = h->err_cell[1] = h->err_cell[2] = 0;
  }
  return code;
}

int hip_fail(evh_ctx *h, hipError_t e) { return fail(h, EVH_ERR_HIP, hipGetErrorString(e)); }

#define HIP_TRY(x)                         \
  do {                                     \
    hipError_t e__ = (x);                  \
    if (e__ != hipSuccess) return hip_fail(h, e__); \
  } while (0)

// ---------------------------------------------------------------- small kernels
__global__ void k_reset_status(Status *st, int keep_next) {
  if (threadIdx.x == 0) {
    st->vmin_bits = (unsigned long long)__double_as_longlong(kVInit);
    if (!keep_next) {
      st->next_code = 0;
      st->next_cell = 0;
    }
  }
}

__global__ void k_clear_code(Status *st) {
  if (threadIdx.x == 0) {
    st->code = 0;
    st->cell = 0;
    st->next_code = 0;
    st->next_cell = 0;
    st->vmin_bits = (unsigned long long)__double_as_longlong(kVInit);
  }
}

// AoS chunk (cells x 6, double) -> SoA storage at cell offset `base`
template <typename S>
__global__ void k_aos_to_soa(const double *__restrict__ src, S *__restrict__ dst, long long cells, long long base,
                             long long n) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < cells; c += (long long)gridDim.x * 256) {
#pragma unroll
    for (int m = 0; m < 6; m++) dst[m * n + base + c] = S(src[6 * c + m]);
  }
}

template <typename S>
__global__ void k_soa_to_aos(const S *__restrict__ src, double *__restrict__ dst, long long cells, long long base,
                             long long n) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < cells; c += (long long)gridDim.x * 256) {
#pragma unroll
    for (int m = 0; m < 6; m++) dst[6 * c + m] = double(src[m * n + base + c]);
  }
}

// Compensated sums of q1, q5, q6 (snapshot.f90:169-172).  Every addition is
// an error-free transformation (TwoSum: s + x = t + err exactly), so each
// thread, the block tree and the host carry a (sum, error) pair and the
// result is the exact sum rounded once, up to the error of adding the error
// terms (relative 1e-16 of the magnitudes for n up to 2^40 cells).
constexpr int kSumBlocks = 1024;

__device__ __host__ __forceinline__ void two_sum(double &s, double &c, double x) {
  const double t = s + x;
  const double z = t - s;
  c += (s - (t - z)) + (x - z);
  s = t;
}

template <typename S>
__global__ void __launch_bounds__(256) k_sums(const S *__restrict__ q, long long n, double *partials) {
  __shared__ double sh[6][256];
  double s[3] = {0, 0, 0}, comp[3] = {0, 0, 0};
  const int comps[3] = {0, 4, 5};
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
#pragma unroll
    for (int j = 0; j < 3; j++) two_sum(s[j], comp[j], double(q[comps[j] * n + c]));
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    sh[j][threadIdx.x] = s[j];
    sh[3 + j][threadIdx.x] = comp[j];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double a = sh[j][threadIdx.x], e = sh[3 + j][threadIdx.x] + sh[3 + j][threadIdx.x + w];
        two_sum(a, e, sh[j][threadIdx.x + w]);
        sh[j][threadIdx.x] = a;
        sh[3 + j][threadIdx.x] = e;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 6; j++) partials[6 * blockIdx.x + j] = sh[j][0];
}

int grid_of(long long work) {
  long long b = (work + 255) / 256;
  if (b > 16384) b = 16384;
  return b < 1 ? 1 : (int)b;
}

bool mixed(const evh_ctx *h) { return h->cfg.precision == EVH_FP32_MIXED; }

void accumulate_timing(evh_ctx *h) {
  for (int i = 0; i < kStages; i++) {
    if (!h->tev_pending[i]) continue;
    float ms = 0.f;
    if (hipEventSynchronize(h->tev[i][1]) == hipSuccess &&
        hipEventElapsedTime(&ms, h->tev[i][0], h->tev[i][1]) == hipSuccess) {
      h->stage_ms[i] += ms;
      h->stage_launches[i] += 1;
    }
    h->tev_pending[i] = false;
  }
}

// events around one launch of a stage, on the context's stream
hipError_t timer_begin(evh_ctx *h, int stage) {
  accumulate_timing(h);
  return h->timing ? hipEventRecord(h->tev[stage][0], h->stream) : hipSuccess;
}
hipError_t timer_end(evh_ctx *h, int stage) {
  if (!h->timing) return hipSuccess;
  const hipError_t e = hipEventRecord(h->tev[stage][1], h->stream);
  if (e == hipSuccess) h->tev_pending[stage] = true;
  return e;
}

// Sync and surface device-side errors.  Errors stay sticky until read here.
int check_status(evh_ctx *h) {
  // the status block is read behind the queued work (pinned host copy, one
  // wait instead of a wait and a blocking copy)
  HIP_TRY(hipMemcpyAsync(h->st_host, h->st, sizeof(Status), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  accumulate_timing(h);
  if (h->st_host->code) {
    const int code = h->st_host->code;
    const long long cell = (long long)h->st_host->cell;
    hipLaunchKernelGGL(k_clear_code, dim3(1), dim3(64), 0, h->stream, h->st);
    hipStreamSynchronize(h->stream);
    h->v_valid = false;
    return fail(h, code, "device", cell);
  }
  return EVH_OK;
}

int ensure_staged_arrays(evh_ctx *h) {
  if (h->w) return EVH_OK;
  const Grid &G = h->grid;
  HIP_TRY(hipMalloc(&h->w, sizeof(double) * 5 * G.nw()));
  HIP_TRY(hipMalloc(&h->evo, sizeof(double) * 5 * G.ne()));
  HIP_TRY(hipMalloc(&h->s, sizeof(double) * 15 * G.ne()));
  HIP_TRY(hipMemsetAsync(h->w, 0, sizeof(double) * 5 * G.nw(), h->stream));
  HIP_TRY(hipMemsetAsync(h->evo, 0, sizeof(double) * 5 * G.ne(), h->stream));  // main.f90:76-77
  HIP_TRY(hipMemsetAsync(h->s, 0, sizeof(double) * 15 * G.ne(), h->stream));
  h->dev_bytes += (long long)sizeof(double) * (5 * G.nw() + 20 * G.ne());
  return EVH_OK;
}

int validate(const evh_config *c) {
  if (c->nx < 1 || c->ny < 1 || c->nz < 1) return EVH_ERR_CONFIG;
  if (c->solver != 1 && c->solver != 2) return EVH_ERR_SOLVER;
  if (c->limiter < 1 || c->limiter > 3) return EVH_ERR_LIMITER;
  if (c->boundtype < 1 || c->boundtype > 3) return EVH_ERR_CONFIG;
  // degenerate dimensions the reference handles: 1-D (ny=nz=1), 2-D (nz=1)
  const bool oned = c->ny == 1 && c->nz == 1, twod = c->nz == 1;
  if ((bool)c->onedimension != oned || (bool)c->twodimension != twod) return EVH_ERR_CONFIG;
  if (c->nx < 2 || (!oned && c->ny < 2)) return EVH_ERR_CONFIG;
  if (c->precision != EVH_FP64 && c->precision != EVH_FP32_MIXED) return EVH_ERR_CONFIG;
  if (c->precision == EVH_FP32_MIXED && !fused_supported(*c)) return EVH_ERR_CONFIG;
  // conduction: simplekappa only (the reference's other branch reads an
  // unassigned dtempdx, conduction.f90:134-141), no dual energy (thermalstep
  // evaluates pressure on q a second time, timestep.f90:87); degenerate
  // directions are skipped as in conduction.f90:46,72,98
  if (c->conduction && (!c->simplekappa || c->dualenergy || !(c->temp0 > 0.0))) return EVH_ERR_CONFIG;
  if (c->conduction && c->precision == EVH_FP32_MIXED && !fused_supported(*c)) return EVH_ERR_CONFIG;
  // a run without hydro (main.f90:119,155) has only conduction to do
  if (c->hydro_off && !c->conduction) return EVH_ERR_CONFIG;
  // gravity: the FFT Poisson solve of gravitation.f90 on one rank, on 3-D,
  // 2-D (nz = 1) and 1-D (ny = nz = 1) grids, the reference's reduced
  // dimensions (parameter.f90:173-182); x always has two cells or more
  if (c->gravity && (c->nx < 2 || !(c->h2 > 0.0) || c->nyqx < 1 || c->nyqy < 1 || c->nyqz < 1 ||
                     (c->supercomoving && !(c->fb > 0.0))))
    return EVH_ERR_CONFIG;
  return EVH_OK;
}

}  // namespace

namespace evh {
KParams make_kparams(const evh_config &c) {
  KParams k;
  k.nx = c.nx; k.ny = c.ny; k.nz = c.nz;
  k.boundtype = c.boundtype; k.solver = c.solver; k.limiter = c.limiter;
  k.secondorder = c.secondorder; k.dualenergy = c.dualenergy;
  k.onedimension = c.onedimension; k.twodimension = c.twodimension;
  k.g = c.g; k.g1 = c.g1; k.g1inv = c.g1inv; k.gg1 = c.gg1; k.cfl = c.cfl;
  k.de_thresh1 = c.de_thresh1;
  k.dx = c.dx; k.dy = c.dy; k.dz = c.dz;
  k.dxinv = c.dxinv; k.dyinv = c.dyinv; k.dzinv = c.dzinv;
  k.mach_fp64 = c.mach_fp64;
  // Mach M  <=>  kin/e = g (g-1) M^2 / 2  <=>  kin/E = R / (1 + R)
  const double R = 0.5 * c.g * c.g1 * c.mach_fp64 * c.mach_fp64;
  k.kin_frac64 = c.mach_fp64 > 0 ? R / (1.0 + R) : -1.0;
  const double delta = 2.0 / 3.0 - c.g1;
  k.g1_third = std::fabs(delta) <= 1e-6 ? 1 : 0;
  k.g1_dln2 = delta * 0.69314718055994530942;
  k.conduction = c.conduction;
  k.c_cond = c.c_cond;
  k.temp0 = c.temp0;
  k.kappa0 = c.kappa0;
  return k;
}

hipError_t launch_reset_status(evh_ctx *h, bool keep_next) {
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, h->stream, h->st, keep_next ? 1 : 0);
  return hipGetLastError();
}
}  // namespace evh

extern "C" {

int evh_create(const evh_config *cfg, evh_ctx **out) {
  if (!cfg || !out) return EVH_ERR_CONFIG;
  *out = nullptr;
  const int v = validate(cfg);
  if (v) return v;
  evh_ctx *h = new (std::nothrow) evh_ctx();
  if (!h) return EVH_ERR_HIP;
  h->cfg = *cfg;
  h->kp = make_kparams(*cfg);
  h->grid = Grid{cfg->nx, cfg->ny, cfg->nz};
  h->qelem = mixed(h) ? sizeof(float) : sizeof(double);
  h->mode = fused_supported(*cfg) ? EVH_MODE_FUSED : EVH_MODE_STAGED;
  // launch shape measured fastest on MI355X for the 512^3 blast (16x12 tiles,
  // 64-plane chunks, the higher waves-per-SIMD budget: FP64 runs at 2
  // waves/SIMD, FP32 at 4)
  h->tile_ty = 12;
  h->zchunk = 0;  // automatic: 64 planes, fewer on small grids
  h->wpe_hi = 1;

  auto bail = [&](hipError_t e) {
    int r = hip_fail(h, e);
    evh_destroy(h);
    return r;
  };
  hipError_t e;
  if ((e = hipGetDevice(&h->device)) != hipSuccess) return bail(e);
  if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e);
  for (auto &ev : h->events)
    if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(e);
  for (auto &pair : h->tev)
    for (auto &ev : pair)
      if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(e);
  const long long n = h->grid.n();
  for (int b = 0; b < 2; b++) {
    if ((e = hipMalloc(&h->qbuf[b], h->qelem * 6 * n)) != hipSuccess) return bail(e);
    h->dev_bytes += (long long)h->qelem * 6 * n;
  }
  if (h->cfg.conduction) {
    if ((e = hipMalloc(&h->tbuf, h->qelem * n)) != hipSuccess) return bail(e);
    h->dev_bytes += (long long)h->qelem * n;
  }
  if (h->cfg.gravity && (e = gravity_init(h)) != hipSuccess) return bail(e);
  if ((e = h->xfer.init(n)) != hipSuccess) return bail(e);
  if ((e = hipMalloc(&h->st, sizeof(Status))) != hipSuccess) return bail(e);
  if ((e = hipHostMalloc(&h->st_host, sizeof(Status))) != hipSuccess) return bail(e);
  if ((e = hipMalloc(&h->partials, sizeof(double) * 6 * kSumBlocks)) != hipSuccess) return bail(e);
  h->dev_bytes += h->xfer.device_bytes() + (long long)sizeof(Status) + (long long)sizeof(double) * 6 * kSumBlocks;
  hipLaunchKernelGGL(k_clear_code, dim3(1), dim3(64), 0, h->stream, h->st);
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bail(e);
  *out = h;
  return EVH_OK;
}

int evh_destroy(evh_ctx *h) {
  if (!h) return EVH_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto p : h->qbuf) if (p) hipFree(p);
  if (h->tbuf) hipFree(h->tbuf);
  gravity_release(h);
  if (h->w) hipFree(h->w);
  if (h->evo) hipFree(h->evo);
  if (h->s) hipFree(h->s);
  h->xfer.release();
  if (h->st) hipFree(h->st);
  if (h->st_host) hipHostFree(h->st_host);
  if (h->partials) hipFree(h->partials);
  for (auto p : h->zhalo) if (p) hipFree(p);
  for (auto p : h->thalo) if (p) hipFree(p);
  for (auto ev : h->events) if (ev) hipEventDestroy(ev);
  for (auto &pair : h->tev)
    for (auto ev : pair)
      if (ev) hipEventDestroy(ev);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return EVH_OK;
}

int evh_set_mode(evh_ctx *h, int mode) {
  if (!h) return EVH_ERR_STATE;
  if (mode != EVH_MODE_FUSED && mode != EVH_MODE_STAGED && mode != EVH_MODE_FUSED_EXACT) return fail(h, EVH_ERR_CONFIG);
  if (mode != EVH_MODE_STAGED && !fused_supported(h->cfg))
    return fail(h, EVH_ERR_CONFIG, "fused needs a 2-D or 3-D grid");
  if (mode != EVH_MODE_FUSED && mixed(h)) return fail(h, EVH_ERR_CONFIG, "staged and exact kernels are FP64");
  if (mode == EVH_MODE_STAGED && (h->zslab[0] || h->zslab[1]))
    return fail(h, EVH_ERR_CONFIG, "a z-slab runs the fused kernels only");
  h->mode = mode;
  h->v_valid = h->tv_valid = h->t_valid = false;  // limits and T are formed in the mode's arithmetic
  h->w_valid = false;
  return EVH_OK;
}

int evh_upload_q(evh_ctx *h, const double *q) {
  if (!h || !q) return EVH_ERR_STATE;
  const long long n = h->grid.n();
  void *dst = h->qbuf[h->cur];
  const bool f32 = mixed(h);
  hipStream_t s = h->stream;
  HIP_TRY(h->xfer.upload(q, n, s, [=](const double *stage, long long cells, long long base) {
    if (f32)
      hipLaunchKernelGGL(k_aos_to_soa<float>, dim3(grid_of(cells)), dim3(256), 0, s, stage, (float *)dst, cells, base, n);
    else
      hipLaunchKernelGGL(k_aos_to_soa<double>, dim3(grid_of(cells)), dim3(256), 0, s, stage, (double *)dst, cells,
                         base, n);
    return hipGetLastError();
  }));
  h->v_valid = false;
  h->tv_valid = false;
  h->t_valid = false;
  h->w_valid = false;
  h->rho_sums_valid = false;
  return EVH_OK;
}

int evh_download_q(evh_ctx *h, double *q) {
  if (!h || !q) return EVH_ERR_STATE;
  int r = check_status(h);
  if (r) return r;
  const long long n = h->grid.n();
  const void *src = h->qbuf[h->cur];
  const bool f32 = mixed(h);
  hipStream_t s = h->stream;
  HIP_TRY(h->xfer.download(q, n, s, [=](double *stage, long long cells, long long base) {
    if (f32)
      hipLaunchKernelGGL(k_soa_to_aos<float>, dim3(grid_of(cells)), dim3(256), 0, s, (const float *)src, stage, cells,
                         base, n);
    else
      hipLaunchKernelGGL(k_soa_to_aos<double>, dim3(grid_of(cells)), dim3(256), 0, s, (const double *)src, stage,
                         cells, base, n);
    return hipGetLastError();
  }));
  return EVH_OK;
}

// ---------------------------------------------------------------- step-level routines
int evh_primitive(evh_ctx *h) {
  if (!h) return EVH_ERR_STATE;
  if (h->mode == EVH_MODE_STAGED) {
    int r = ensure_staged_arrays(h);
    if (r) return r;
    HIP_TRY(launch_primitive(h));
    h->w_valid = true;
    h->bound_done = false;
    return check_status(h);
  }
  if (!h->v_valid) {
    HIP_TRY(launch_reset_status(h, false));
    HIP_TRY(launch_cfl(h));
    h->v_valid = true;
  }
  return EVH_OK;
}

int evh_bound(evh_ctx *h) {
  if (!h) return EVH_ERR_STATE;
  if (h->mode == EVH_MODE_STAGED) {
    if (!h->w_valid) return fail(h, EVH_ERR_STATE, "bound before primitive");
    HIP_TRY(launch_bound(h));
    h->bound_done = true;
  }
  return EVH_OK;
}

int evh_hydrostep(evh_ctx *h, double *dt) {
  if (!h || !dt) return EVH_ERR_STATE;
  if (h->mode == EVH_MODE_STAGED) {
    if (!h->w_valid) return fail(h, EVH_ERR_STATE, "hydrostep before primitive");
    HIP_TRY(launch_reset_status(h, false));
    HIP_TRY(launch_hydrostep(h));
  } else if (!h->v_valid) {
    return fail(h, EVH_ERR_STATE, "hydrostep before primitive");
  }
  int r = check_status(h);
  if (r) return r;
  if (h->st_host->next_code) {
    h->v_valid = false;
    return fail(h, (int)h->st_host->next_code, "primitive", (long long)h->st_host->next_cell);
  }
  double v;
  const unsigned long long bits = h->st_host->vmin_bits;
  std::memcpy(&v, &bits, sizeof v);
  const double c = h->cfg.cfl * v;
  *dt = c < *dt ? c : *dt;  // timestep.f90:49
  if (*dt <= 0.0) return fail(h, EVH_ERR_DT);
  return EVH_OK;
}

int evh_compevo(evh_ctx *h, double dt) {
  if (!h) return EVH_ERR_STATE;
  if (h->mode == EVH_MODE_STAGED) {
    if (!h->bound_done) return fail(h, EVH_ERR_STATE, "compevo before bound");
    HIP_TRY(launch_compevo(h, dt));
  }
  h->dt_evo = dt;
  h->evo_valid = true;
  return EVH_OK;
}

int evh_sweep(evh_ctx *h, double dt) {
  if (!h) return EVH_ERR_STATE;
  if (h->cfg.secondorder && (!h->evo_valid || h->dt_evo != dt))
    return fail(h, EVH_ERR_STATE, "sweep without compevo at the same dt");
  h->evo_valid = false;
  if (h->mode == EVH_MODE_STAGED) {
    if (!h->bound_done) return fail(h, EVH_ERR_STATE, "sweep before bound");
    HIP_TRY(timer_begin(h, EVH_STAGE_SWEEP));
    HIP_TRY(launch_sweep(h, dt));
    HIP_TRY(timer_end(h, EVH_STAGE_SWEEP));
    h->w_valid = false;
    h->bound_done = false;
    h->tv_valid = false;
    h->t_valid = false;
    return check_status(h);
  }
  if (!h->v_valid) return fail(h, EVH_ERR_STATE, "sweep before primitive");
  HIP_TRY(launch_reset_status(h, false));
  HIP_TRY(timer_begin(h, EVH_STAGE_SWEEP));
  HIP_TRY(launch_fused_step(h, dt));
  HIP_TRY(timer_end(h, EVH_STAGE_SWEEP));
  h->cur = 1 - h->cur;
  h->v_valid = true;
  h->tv_valid = false;
  h->t_valid = h->tbuf != nullptr;  // the sweep stored T of the new state
  return EVH_OK;
}

// timestep.f90:72-104: dt = min(c_cond dx^2 min(rho/kappa) / temp0, dt)
int evh_thermalstep(evh_ctx *h, double *dt) {
  if (!h || !dt) return EVH_ERR_STATE;
  if (!h->cfg.conduction) return fail(h, EVH_ERR_STATE, "thermalstep without conduction");
  if (!h->tv_valid) {
    HIP_TRY(launch_thermal_min(h));
    h->tv_valid = true;
  }
  int r = check_status(h);
  if (r) return r;
  if (h->st_host->next_code) {
    h->v_valid = h->tv_valid = false;
    return fail(h, (int)h->st_host->next_code, "thermalstep", (long long)h->st_host->next_cell);
  }
  double v;
  const unsigned long long bits = h->st_host->tvmin_bits;
  std::memcpy(&v, &bits, sizeof v);
  v = h->cfg.c_cond * (h->cfg.dx * h->cfg.dx) * v / h->cfg.temp0;
  *dt = v < *dt ? v : *dt;
  if (*dt <= 0.0) return fail(h, EVH_ERR_DT);
  return EVH_OK;
}

// conduction.f90:26-123 on the state after the sweep, with the primitive and
// bound of main.f90:171-172 folded into the kernel
int evh_thermalconduction(evh_ctx *h, double dt) {
  if (!h) return EVH_ERR_STATE;
  if (!h->cfg.conduction) return fail(h, EVH_ERR_STATE, "thermalconduction without conduction");
  if (!h->t_valid) HIP_TRY(launch_temp(h));
  HIP_TRY(launch_reset_status(h, false));
  HIP_TRY(timer_begin(h, EVH_STAGE_CONDUCTION));
  HIP_TRY(launch_conduction(h, dt));
  HIP_TRY(timer_end(h, EVH_STAGE_CONDUCTION));
  h->v_valid = true;
  h->tv_valid = true;
  h->t_valid = false;  // E changed
  h->w_valid = false;
  h->bound_done = false;
  return EVH_OK;
}

// gravitation.f90:26-52: the potential of the current q (phi stays on the device)
int evh_gravpotential(evh_ctx *h, double cosmo) {
  if (!h) return EVH_ERR_STATE;
  if (!h->cfg.gravity) return fail(h, EVH_ERR_STATE, "gravpotential without gravity");
  h->cosmo = cosmo;
  HIP_TRY(timer_begin(h, EVH_STAGE_GRAVPOTENTIAL));
  HIP_TRY(launch_gravpotential(h, cosmo));
  HIP_TRY(timer_end(h, EVH_STAGE_GRAVPOTENTIAL));
  h->phi_valid = true;
  return EVH_OK;
}

// gravitation.f90:172-254 with the last potential
int evh_gravintegration(evh_ctx *h, double dt) {
  if (!h) return EVH_ERR_STATE;
  if (!h->cfg.gravity) return fail(h, EVH_ERR_STATE, "gravintegration without gravity");
  if (!h->phi_valid) return fail(h, EVH_ERR_STATE, "gravintegration before gravpotential");
  HIP_TRY(launch_reset_status(h, false));
  HIP_TRY(timer_begin(h, EVH_STAGE_GRAVINTEGRATION));
  HIP_TRY(launch_gravintegration(h, dt));
  HIP_TRY(timer_end(h, EVH_STAGE_GRAVINTEGRATION));
  h->v_valid = true;                // the kernel formed the CFL minimum of the new state
  h->tv_valid = false;
  h->t_valid = h->tbuf != nullptr;  // and T, with conduction on
  h->w_valid = false;
  h->bound_done = false;
  return EVH_OK;
}

int evh_download_phi(evh_ctx *h, double *phi) {
  if (!h || !phi) return EVH_ERR_STATE;
  if (!h->phi_valid) return fail(h, EVH_ERR_STATE, "no potential yet");
  int r = check_status(h);
  if (r) return r;
  HIP_TRY(hipMemcpy(phi, h->phi, sizeof(double) * h->grid.n(), hipMemcpyDeviceToHost));
  return EVH_OK;
}

int evh_step(evh_ctx *h, double *t, double tend, double *dt, int *go) {
  if (!h || !t || !dt || !go) return EVH_ERR_STATE;
  if (h->zslab[0] || h->zslab[1])
    return fail(h, EVH_ERR_STATE, "a z-slab steps through hydrostep and sweep, with the exchange between");
  int r;
  if ((r = evh_primitive(h))) return r;
  if ((r = evh_bound(h))) return r;
  *dt = 1e99;  // main.f90:118
  const bool hydro = !h->cfg.hydro_off;
  if (hydro && (r = evh_hydrostep(h, dt))) return r;                 // main.f90:119
  if (h->cfg.conduction && (r = evh_thermalstep(h, dt))) return r;  // main.f90:121
  if (tend - *t <= *dt) {  // timestep.f90:163-166
    *dt = tend - *t;
    *go = 0;
  }
  *t = *t + *dt;
  if (hydro) {  // compevo feeds only the sweep (main.f90:151,155)
    if (h->cfg.secondorder && (r = evh_compevo(h, *dt))) return r;
    if ((r = evh_sweep(h, *dt))) return r;
  }
  if (h->cfg.gravity) {  // main.f90:157-164
    if ((r = evh_gravpotential(h, h->cosmo))) return r;
    if ((r = evh_gravintegration(h, *dt))) return r;
  }
  if (h->cfg.conduction) return evh_thermalconduction(h, *dt);  // main.f90:170-175
  return EVH_OK;
}

int evh_run(evh_ctx *h, int nsteps, double *t, double tend, double *dt, int *done) {
  if (!h || !t || !dt || !done) return EVH_ERR_STATE;
  int go = 1;
  *done = 0;
  for (int k = 0; k < nsteps && go; k++) {
    int r = evh_step(h, t, tend, dt, &go);
    if (r) return r;
    *done = k + 1;
  }
  return check_status(h);
}

int evh_sums(evh_ctx *h, double out[3]) {
  if (!h || !out) return EVH_ERR_STATE;
  int r = check_status(h);
  if (r) return r;
  const long long n = h->grid.n();
  if (mixed(h))
    hipLaunchKernelGGL(k_sums<float>, dim3(kSumBlocks), dim3(256), 0, h->stream, (const float *)h->qbuf[h->cur], n,
                       h->partials);
  else
    hipLaunchKernelGGL(k_sums<double>, dim3(kSumBlocks), dim3(256), 0, h->stream, (const double *)h->qbuf[h->cur], n,
                       h->partials);
  HIP_TRY(hipGetLastError());
  double part[6 * kSumBlocks];
  HIP_TRY(hipMemcpyAsync(part, h->partials, sizeof part, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int j = 0; j < 3; j++) {
    double s = 0, c = 0;
    for (int b = 0; b < kSumBlocks; b++) {
      two_sum(s, c, part[6 * b + j]);
      c += part[6 * b + 3 + j];
    }
    out[j] = s + c;
  }
  return EVH_OK;
}

// ---------------------------------------------------------------- staged-array views
int evh_download_w(evh_ctx *h, double *w) {
  if (!h || !w) return EVH_ERR_STATE;
  if (h->mode != EVH_MODE_STAGED || !h->w) return fail(h, EVH_ERR_STATE, "w exists in staged mode only");
  int r = check_status(h);
  if (r) return r;
  const Grid &G = h->grid;
  const long long nw = G.nw();
  double *tmp = new (std::nothrow) double[5 * nw];
  if (!tmp) return fail(h, EVH_ERR_HIP, "host allocation");
  hipError_t e = hipMemcpy(tmp, h->w, sizeof(double) * 5 * nw, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (long long c = 0; c < nw; c++)
      for (int m = 0; m < 5; m++) w[5 * c + m] = tmp[m * nw + c];
  delete[] tmp;
  return e == hipSuccess ? EVH_OK : hip_fail(h, e);
}

int evh_download_evo_s(evh_ctx *h, double *evo, double *s) {
  if (!h || !evo || !s) return EVH_ERR_STATE;
  if (h->mode != EVH_MODE_STAGED || !h->evo) return fail(h, EVH_ERR_STATE, "evo/s exist in staged mode only");
  int r = check_status(h);
  if (r) return r;
  const long long ne = h->grid.ne();
  double *tmp = new (std::nothrow) double[15 * ne];
  if (!tmp) return fail(h, EVH_ERR_HIP, "host allocation");
  hipError_t e = hipMemcpy(tmp, h->evo, sizeof(double) * 5 * ne, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (long long c = 0; c < ne; c++)
      for (int m = 0; m < 5; m++) evo[5 * c + m] = tmp[m * ne + c];
  if (e == hipSuccess) e = hipMemcpy(tmp, h->s, sizeof(double) * 15 * ne, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (long long c = 0; c < ne; c++)
      for (int d = 0; d < 3; d++)
        for (int m = 0; m < 5; m++) s[15 * c + 5 * d + m] = tmp[(m + 5 * d) * ne + c];
  delete[] tmp;
  return e == hipSuccess ? EVH_OK : hip_fail(h, e);
}

int evh_last_error(evh_ctx *h, int *code, int cell[3], char *msg, int msglen) {
  if (!h) return EVH_ERR_STATE;
  if (code) *code = h->err_code;
  if (cell) for (int i = 0; i < 3; i++) cell[i] = h->err_cell[i];
  if (msg && msglen > 0) {
    std::strncpy(msg, h->err_msg.c_str(), (size_t)msglen - 1);
    msg[msglen - 1] = 0;
  }
  return EVH_OK;
}

int evh_event_record(evh_ctx *h, int slot) {
  if (!h || slot < 0 || slot >= 8) return EVH_ERR_STATE;
  HIP_TRY(hipEventRecord(h->events[slot], h->stream));
  return EVH_OK;
}

int evh_event_elapsed(evh_ctx *h, int a, int b, float *ms) {
  if (!h || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return EVH_ERR_STATE;
  HIP_TRY(hipEventSynchronize(h->events[b]));
  HIP_TRY(hipEventElapsedTime(ms, h->events[a], h->events[b]));
  return EVH_OK;
}

int evh_synchronize(evh_ctx *h) {
  if (!h) return EVH_ERR_STATE;
  return check_status(h);
}

int evh_set_tuning(evh_ctx *h, int tile_ty, int wpe_hi, int zchunk) {
  if (!h) return EVH_ERR_STATE;
  if ((tile_ty != 8 && tile_ty != 12) || (wpe_hi != 0 && wpe_hi != 1) || zchunk < 0)
    return fail(h, EVH_ERR_CONFIG, "tile_ty must be 8 or 12, wpe_hi 0 or 1, zchunk >= 0");
  h->tile_ty = tile_ty;
  h->wpe_hi = wpe_hi != 0;
  h->zchunk = zchunk;
  return EVH_OK;
}

int evh_get_tuning(evh_ctx *h, int *tile_ty, int *wpe_hi, int *zchunk) {
  if (!h || !tile_ty || !wpe_hi || !zchunk) return EVH_ERR_STATE;
  *tile_ty = h->tile_ty;
  *wpe_hi = h->wpe_hi ? 1 : 0;
  *zchunk = h->zchunk;
  return EVH_OK;
}

int evh_set_timing(evh_ctx *h, int on) {
  if (!h) return EVH_ERR_STATE;
  int r = check_status(h);
  if (r) return r;
  h->timing = on != 0;
  for (int i = 0; i < kStages; i++) {
    h->stage_ms[i] = 0.0;
    h->stage_launches[i] = 0;
  }
  return EVH_OK;
}

int evh_kernel_stats(evh_ctx *h, double *total_ms, long long *launches) {
  if (!h || !total_ms || !launches) return EVH_ERR_STATE;
  int r = check_status(h);
  if (r) return r;
  *total_ms = h->stage_ms[EVH_STAGE_SWEEP];
  *launches = h->stage_launches[EVH_STAGE_SWEEP];
  return EVH_OK;
}

int evh_conduction_stats(evh_ctx *h, double *total_ms, long long *launches) {
  if (!h || !total_ms || !launches) return EVH_ERR_STATE;
  int r = check_status(h);
  if (r) return r;
  *total_ms = h->stage_ms[EVH_STAGE_CONDUCTION];
  *launches = h->stage_launches[EVH_STAGE_CONDUCTION];
  return EVH_OK;
}

int evh_stage_stats(evh_ctx *h, int stage, double *total_ms, long long *launches) {
  if (!h || !total_ms || !launches) return EVH_ERR_STATE;
  if (stage < 0 || stage >= kStages) return fail(h, EVH_ERR_CONFIG, "no such stage");
  int r = check_status(h);
  if (r) return r;
  *total_ms = h->stage_ms[stage];
  *launches = h->stage_launches[stage];
  return EVH_OK;
}

// ---------------------------------------------------------------- z-slabs
// One rank's share of a grid decomposed into z-slabs (parallel.f90:26-66):
// the ghost planes on a side with a neighbour are the neighbour's planes
// (communicate, parallel.f90:317-381), which the caller places in the halo
// buffer of that side; a side without one keeps the z rules of bound
// (boundary.f90:90-110).  q, not w, crosses: the march forms the primitive
// of every plane it reads.
int evh_slab(evh_ctx *h, int lo, int hi, void **lo_halo, void **hi_halo) {
  if (!h) return EVH_ERR_STATE;
  if (h->mode == EVH_MODE_STAGED) return fail(h, EVH_ERR_CONFIG, "a z-slab runs the fused kernels only");
  if (h->grid.nz < 2) return fail(h, EVH_ERR_CONFIG, "a z-slab has two planes or more");
  if (h->cfg.gravity) return fail(h, EVH_ERR_CONFIG, "a z-slab has no distributed Poisson solve");
  const bool on[2] = {lo != 0, hi != 0};
  const size_t pl = (size_t)h->grid.nx * h->grid.ny;
  for (int side = 0; side < 2; side++) {
    if (on[side] && !h->zhalo[side]) {
      HIP_TRY(hipMalloc(&h->zhalo[side], 12 * pl * h->qelem));
      h->dev_bytes += (long long)(12 * pl * h->qelem);
      HIP_TRY(hipMemsetAsync(h->zhalo[side], 0, 12 * pl * h->qelem, h->stream));
    }
    if (on[side] && h->cfg.conduction && !h->thalo[side]) {
      HIP_TRY(hipMalloc(&h->thalo[side], pl * h->qelem));
      h->dev_bytes += (long long)(pl * h->qelem);
      HIP_TRY(hipMemsetAsync(h->thalo[side], 0, pl * h->qelem, h->stream));
    }
    h->zslab[side] = on[side];
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (lo_halo) *lo_halo = on[0] ? h->zhalo[0] : nullptr;
  if (hi_halo) *hi_halo = on[1] ? h->zhalo[1] : nullptr;
  return EVH_OK;
}

// The two lowest and two highest planes of the current q, [6][2][ny][nx],
// into device memory: the planes communicate sends to the neighbours below
// and above.
int evh_slab_faces(evh_ctx *h, void *lo_dst, void *hi_dst) {
  if (!h) return EVH_ERR_STATE;
  if (h->grid.nz < 2) return fail(h, EVH_ERR_CONFIG, "a z-slab has two planes or more");
  const size_t pl = (size_t)h->grid.nx * h->grid.ny, e = h->qelem;
  const size_t n = (size_t)h->grid.n();
  const char *q = (const char *)h->qbuf[h->cur];
  if (lo_dst)
    HIP_TRY(hipMemcpy2DAsync(lo_dst, 2 * pl * e, q, n * e, 2 * pl * e, 6, hipMemcpyDeviceToDevice, h->stream));
  if (hi_dst)
    HIP_TRY(hipMemcpy2DAsync(hi_dst, 2 * pl * e, q + (h->grid.nz - 2) * pl * e, n * e, 2 * pl * e, 6,
                             hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVH_OK;
}

// With conduction on: the halo buffers of T for the plane next to the slab
// on each side with a neighbour ([ny][nx], the storage type), filled before
// each evh_thermalconduction (main.f90:170-175 communicates w after the
// sweep); and T of the slab's lowest and highest planes, the ones its
// neighbours need, into device memory.
int evh_slab_temp(evh_ctx *h, void **lo_halo, void **hi_halo) {
  if (!h) return EVH_ERR_STATE;
  if (!h->cfg.conduction) return fail(h, EVH_ERR_STATE, "slab temperatures without conduction");
  if (lo_halo) *lo_halo = h->zslab[0] ? h->thalo[0] : nullptr;
  if (hi_halo) *hi_halo = h->zslab[1] ? h->thalo[1] : nullptr;
  return EVH_OK;
}

int evh_slab_temp_faces(evh_ctx *h, void *lo_dst, void *hi_dst) {
  if (!h) return EVH_ERR_STATE;
  if (!h->cfg.conduction) return fail(h, EVH_ERR_STATE, "slab temperatures without conduction");
  if (!h->t_valid) {  // T of the current state (after an upload, or a sweep of the staged kernels)
    HIP_TRY(launch_temp(h));
    h->t_valid = true;
  }
  const size_t bytes = (size_t)h->grid.nx * h->grid.ny * h->qelem;
  const char *t = (const char *)h->tbuf;
  if (lo_dst) HIP_TRY(hipMemcpyAsync(lo_dst, t, bytes, hipMemcpyDeviceToDevice, h->stream));
  if (hi_dst)
    HIP_TRY(hipMemcpyAsync(hi_dst, t + (h->grid.nz - 1) * bytes, bytes, hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVH_OK;
}

long long evh_device_bytes(evh_ctx *h) { return h ? h->dev_bytes : 0; }

}  // extern "C"
