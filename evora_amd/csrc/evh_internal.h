// evh_internal.h -- context and kernel launchers of libevora_hip (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/evora_hip.h"
#include "physics.h"
#include "tiling.h"
#include "transfer.h"

namespace evh {

// Device SoA layouts (x fastest):
//   q   [6][nz][ny][nx]                      storage type per precision
//   w   [5][nz+4][ny+4][nx+4]  (staged)      ghost cell g at g+2
//   evo [5][nz+2][ny+2][nx+2]  (staged)      cell c at c (reference index 0..n+1)
//   s   [15][nz+2][ny+2][nx+2] (staged)      component c of direction d at c + 5*d
struct Grid {
  int nx, ny, nz;
  __host__ __device__ long long n() const { return (long long)nx * ny * nz; }
  __host__ __device__ long long nw() const { return (long long)(nx + 4) * (ny + 4) * (nz + 4); }
  __host__ __device__ long long ne() const { return (long long)(nx + 2) * (ny + 2) * (nz + 2); }
};

}  // namespace evh

constexpr int kStages = 4;  // EVH_STAGE_SWEEP .. EVH_STAGE_GRAVINTEGRATION

struct evh_ctx {
  evh_config cfg;
  evh::KParams kp;
  evh::Grid grid;
  int device = 0;
  int mode = EVH_MODE_FUSED;
  int zchunk = 0;   // z planes per fused block (0: default)
  int tile_ty = 12;    // fused tile height: 8 or 12
  bool wpe_hi = true;  // register budget for the higher of the two occupancy targets
  hipStream_t stream = nullptr;
  hipEvent_t events[8] = {};

  // conserved state: two buffers (fused path ping-pongs), storage precision per cfg.precision
  void *qbuf[2] = {nullptr, nullptr};
  int cur = 0;
  size_t qelem = 8;  // bytes per stored value

  // temperature of each cell of the current q (conduction on only): stored
  // by the fused sweep, read by the conduction pass; storage precision
  void *tbuf = nullptr;
  bool t_valid = false;

  // self-gravity (gravity on): potential, its half spectrum (real-to-complex
  // transform along x), the sin^2 tables of the Green's function (x, y, z,
  // full ranges), partial sums of rho and their mean, and the FFT plans
  double *phi = nullptr;
  void *spec = nullptr;
  double *sin2 = nullptr;
  double *rho_part = nullptr;
  double *rho_sums = nullptr;   // the x pass's block totals of rho (power-of-two grids)
  int rho_nsums = 0;
  bool rho_sums_valid = false;  // they sum the state of the last potential, which an upload replaces
  void *plan_fwd = nullptr, *plan_inv = nullptr;  // hipfftHandle
  // power-of-two nz: the z transform is the pencil kernel of gravity.hip and
  // the plans transform x and y (2-D, per z plane)
  bool pencils = false;
  bool xyz_pencils = false;  // power-of-two nx and ny too: no plans, the kernels of gravity.hip transform x and y
  long long spec_plane = 0;  // complex values per z plane of the spectrum
  int spec_pitch = 0;        // complex values per ky row of the spectrum (nx/2+1, padded with pencils)
  void *twiddle = nullptr;   // exp(-2 pi i t / n) for n = nz, then (xyz_pencils) ny and nx
  bool phi_valid = false;
  double cosmo = 1.0;  // expansion factor of the last evh_gravpotential

  // a z-slab of a grid decomposed along z (evh_slab): the ghost planes below
  // ([0]) and above ([1]) from the neighbours' planes in zhalo, [6][2][ny][nx]
  // in the storage type, instead of the z rules of bound
  bool zslab[2] = {false, false};
  void *zhalo[2] = {nullptr, nullptr};
  void *thalo[2] = {nullptr, nullptr};  // with conduction: T of the neighbours' adjacent planes, [ny][nx]

  // staged path arrays (allocated on first use)
  double *w = nullptr, *evo = nullptr, *s = nullptr;

  // host transfers (pinned, double-buffered chunks)
  evh::Xfer xfer;

  // reductions
  evh::Status *st = nullptr;      // device
  evh::Status *st_host = nullptr; // pinned
  double *partials = nullptr;     // per-block sums
  int npartials = 0;

  // step state
  bool v_valid = false;  // st->vmin_bits holds the CFL min of the current q
  bool tv_valid = false; // st->tvmin_bits holds the conduction limit of the current q
  bool w_valid = false;  // staged: w holds primitive of current q (bound applied?)
  bool bound_done = false;
  double dt_evo = 0.0;
  bool evo_valid = false;

  // last error
  int err_code = 0;
  int err_cell[3] = {0, 0, 0};
  std::string err_msg;

  long long dev_bytes = 0;

  // kernel timing (evh_set_timing)
  bool timing = false;
  // per stage (EVH_STAGE_*): start/end events of the last launch, totals
  hipEvent_t tev[kStages][2] = {};
  bool tev_pending[kStages] = {};
  double stage_ms[kStages] = {};
  long long stage_launches[kStages] = {};
};

namespace evh {

// staged kernels (staged.hip)
hipError_t launch_primitive(evh_ctx *h);
hipError_t launch_bound(evh_ctx *h);
hipError_t launch_hydrostep(evh_ctx *h);
hipError_t launch_compevo(evh_ctx *h, double dt);
hipError_t launch_sweep(evh_ctx *h, double dt);

// fused kernels (fused.hip)
hipError_t launch_cfl(evh_ctx *h);                 // primitive + CFL min over cur q
hipError_t launch_fused_step(evh_ctx *h, double dt);  // cur -> other, CFL of the result
hipError_t launch_fused_f64(evh_ctx *h, double dt);
hipError_t launch_fused_f32(evh_ctx *h, double dt);
hipError_t launch_fused_exact(evh_ctx *h, double dt);  // FP64, reference arithmetic
// conduction kernels (conduction.hip)
hipError_t launch_thermal_min(evh_ctx *h);              // conduction limit of cur q
hipError_t launch_temp(evh_ctx *h);                     // tbuf from cur q
hipError_t launch_conduction(evh_ctx *h, double dt);    // cur q in place (tbuf valid), limits of the result
hipError_t launch_thermal_min_exact(evh_ctx *h);        // the same with the reference's arithmetic
hipError_t launch_temp_exact(evh_ctx *h);
hipError_t launch_conduction_exact(evh_ctx *h, double dt);
bool fused_supported(const evh_config &c);

// gravity (gravity.hip)
hipError_t gravity_init(evh_ctx *h);     // buffers, tables and FFT plans
void gravity_release(evh_ctx *h);
hipError_t launch_gravpotential(evh_ctx *h, double cosmo);  // phi of cur q
hipError_t launch_gravintegration(evh_ctx *h, double dt);   // cur q in place with phi, limits (and T) of the result

// KParams of a configuration (api.hip)
KParams make_kparams(const evh_config &c);

// transfers and reductions (api.hip)
hipError_t launch_reset_status(evh_ctx *h, bool keep_next);
hipError_t launch_sums(evh_ctx *h, double *out3);

}  // namespace evh
