// batch.hip -- many independent boxes of uneven sizes in one persistent kernel.
//
// Every box is an independent evora run (own grid, own t and dt, own CFL
// limit).  One step of the whole ensemble is two kernels:
//   k_batch_dt     one thread per box: hydrostep's dt = min(cfl*v, 1d99)
//                  (timestep.f90:26-57) from the box's CFL minimum, then
//                  updatetime's tend clamp (timestep.f90:163-169); stops a box
//                  whose previous step raised an error; resets the work queues.
//   k_batch_fused  persistent blocks take (box, tile) items from 8 per-XCD
//                  work queues (with stealing) and run the tile march of
//                  fused_core.h on them; each box's CFL minimum for the next
//                  step is reduced into its own status block.
// Each box ping-pongs between the two batch buffers on its own: side[b] names
// the buffer that holds its current state, and flips only on a step the box
// takes, so a box stopped by tend or by an error keeps its last state while
// the others run on.  No host synchronisation is needed between steps, and
// the kernels read the sides from device memory, so one step is captured in
// a hipGraph once and replayed.  Launch
// shape (tile height, occupancy target, z chunk, blocks per CU) is picked by
// an autotuner that times candidates on a snapshot of the state and logs them.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fused_core.h"

#ifndef EVH_FUSED_PREC
#define EVH_FUSED_PREC 64
#endif

namespace evh {
namespace batch {

using core::Geo;

struct BoxDev {
  KParams P;
  Geo G;
  long long qoff;  // element offset of the box inside the batch buffers
  int tile0;       // first work-list item of the box
  int pad;
};

struct Queues {
  unsigned next[kXcds];  // items taken from each XCD's queue
  int start[kXcds + 1];  // queue j covers items [start[j], start[j+1])
};

// Per-box state on the device (arrays of nbox)
struct BoxState {
  double *t, *tend, *dt, *cfl;
  int *go, *active, *steps, *err_step;
  int *side;   // buffer (0 or 1) that holds the box's current state
  Status *st;  // per-box status: errors, CFL minimum
};

__device__ __forceinline__ int take_item(Queues *qs, int xcd) {
  for (int i = 0; i < kXcds; i++) {
    const int j = (xcd + i) % kXcds;
    const int len = qs->start[j + 1] - qs->start[j];
    if ((int)__atomic_load_n(&qs->next[j], __ATOMIC_RELAXED) >= len) continue;
    const int idx = (int)atomicAdd(&qs->next[j], 1u);
    if (idx < len) return qs->start[j] + idx;
  }
  return -1;
}

__device__ __forceinline__ int box_of(const BoxDev *boxes, int nbox, int item) {
  int lo = 0, hi = nbox - 1;  // last box with tile0 <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (boxes[mid].tile0 <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, typename S, int SOLVER, int LIMITER, int TY, int WPE>
__global__ void __launch_bounds__(core::TileShape<TY>::NT) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    k_batch_fused(const BoxDev *__restrict__ boxes, int nbox, S *__restrict__ q0, S *__restrict__ q1, BoxState bs,
                  Queues *qs) {
  __shared__ core::Smem<T, TY> sm;
  __shared__ int s_item;
  const int xcd = blockIdx.x % kXcds;
  for (;;) {
    if (threadIdx.x == 0) s_item = take_item(qs, xcd);
    __syncthreads();
    const int item = s_item;
    if (item < 0) break;  // every queue drained: the whole block leaves
    const int b = box_of(boxes, nbox, item);
    if (bs.active[b]) {
      // k_batch_dt has flipped the side of every box that steps: the new
      // state goes to that buffer, the current one is read from the other
      const BoxDev &bx = boxes[b];
      const bool to1 = bs.side[b] != 0;
      core::march<T, S, SOLVER, LIMITER, TY, false>(bx.P, bx.G, to1 ? q0 : q1, to1 ? q1 : q0, bx.qoff, b, bs.dt[b],
                                                    &bs.st[b], &bs.st[b].vmin_bits, item - bx.tile0, sm);
    }
    __syncthreads();  // s_item is rewritten only after every lane has used it
  }
}

#if EVH_FUSED_PREC == 64
__global__ void k_batch_dt(int nbox, BoxState bs, Queues *qs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < kXcds) qs->next[threadIdx.x] = 0;
  if (b >= nbox) return;
  Status &s = bs.st[b];
  bs.active[b] = 0;
  if (!bs.go[b]) return;
  if (s.code || s.next_code) {  // the reference stops the run at an error
    bs.go[b] = 0;
    bs.err_step[b] = bs.steps[b];
    return;
  }
  double v;
  const unsigned long long bits = s.vmin_bits;
  v = __longlong_as_double((long long)bits);
  double dt = bs.cfl[b] * v;
  if (!(dt < 1e99)) dt = 1e99;  // dt = min(cfl*v, 1d99), main.f90:118
  if (dt <= 0.0) {
    s.code = kErrDt;
    bs.go[b] = 0;
    bs.err_step[b] = bs.steps[b];
    return;
  }
  if (bs.tend[b] - bs.t[b] <= dt) {  // timestep.f90:163-166
    dt = bs.tend[b] - bs.t[b];
    bs.go[b] = 0;
  }
  bs.t[b] = bs.t[b] + dt;
  bs.dt[b] = dt;
  bs.steps[b] += 1;
  bs.active[b] = 1;
  bs.side[b] = 1 - bs.side[b];  // this step writes the other buffer
  s.vmin_bits = (unsigned long long)__double_as_longlong(1e6);  // timestep.f90:35
}

// exact primitive + CFL of one box (first step after an upload)
template <typename S>
__global__ void __launch_bounds__(256) k_batch_cfl(KParams P, long long n, const S *__restrict__ q, Status *st) {
  double vmin = 1e6;
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    double qc[6], wc[5];
#pragma unroll
    for (int m = 0; m < 6; m++) qc[m] = double(q[m * n + c]);
    const int e = q2w(P, qc, wc);
    if (e) report_next(st, e, (unsigned long long)c);
    vmin = dmin(vmin, cfl_cell(P, wc));
  }
  for (int off = 32; off > 0; off >>= 1) vmin = dmin(vmin, __shfl_xor(vmin, off));
  if ((threadIdx.x & 63) == 0) atomic_min_pos(&st->vmin_bits, vmin);
}
#endif

}  // namespace batch
}  // namespace evh

namespace evh {
namespace batch {

// launchers for this translation unit's precision
template <typename T, typename S, int SOLVER, int LIMITER, int TY, int WPE>
static void launch(const BoxDev *boxes, int nbox, void *q0, void *q1, const BoxState &bs, Queues *qs, int blocks,
                   hipStream_t stream) {
  hipLaunchKernelGGL((k_batch_fused<T, S, SOLVER, LIMITER, TY, WPE>), dim3(blocks),
                     dim3(core::TileShape<TY>::NT), 0, stream, boxes, nbox, (S *)q0, (S *)q1, bs, qs);
}

template <typename T, typename S, int SOLVER, int LIMITER>
static void launch_shape(int ty, bool hi, const BoxDev *boxes, int nbox, void *q0, void *q1, const BoxState &bs,
                         Queues *qs, int blocks, hipStream_t stream) {
  constexpr int LO = sizeof(T) == 8 ? 2 : 3, HI = LO + 1;
  if (ty == 8) {
    if (hi) launch<T, S, SOLVER, LIMITER, 8, HI>(boxes, nbox, q0, q1, bs, qs, blocks, stream);
    else launch<T, S, SOLVER, LIMITER, 8, LO>(boxes, nbox, q0, q1, bs, qs, blocks, stream);
  } else {
    if (hi) launch<T, S, SOLVER, LIMITER, 12, HI>(boxes, nbox, q0, q1, bs, qs, blocks, stream);
    else launch<T, S, SOLVER, LIMITER, 12, LO>(boxes, nbox, q0, q1, bs, qs, blocks, stream);
  }
}

#if EVH_FUSED_PREC == 32
#define EVH_BATCH_LAUNCH launch_batch_f32
using TT = float;
#else
#define EVH_BATCH_LAUNCH launch_batch_f64
using TT = double;
#endif

hipError_t EVH_BATCH_LAUNCH(int solver, int limiter, int ty, bool hi, const BoxDev *boxes, int nbox, void *q0,
                            void *q1, const BoxState &bs, Queues *qs, int blocks, hipStream_t stream) {
#define EVH_L(SV, LM) launch_shape<TT, TT, SV, LM>(ty, hi, boxes, nbox, q0, q1, bs, qs, blocks, stream)
  if (solver == 1 && limiter == 1) EVH_L(1, 1);
  else if (solver == 1 && limiter == 2) EVH_L(1, 2);
  else if (solver == 1 && limiter == 3) EVH_L(1, 3);
  else if (solver == 2 && limiter == 1) EVH_L(2, 1);
  else if (solver == 2 && limiter == 2) EVH_L(2, 2);
  else if (solver == 2 && limiter == 3) EVH_L(2, 3);
  else return hipErrorInvalidValue;
#undef EVH_L
  return hipGetLastError();
}

}  // namespace batch
}  // namespace evh

// ============================================================================ host side
#if EVH_FUSED_PREC == 64
namespace evh {
namespace batch {
hipError_t launch_batch_f32(int, int, int, bool, const BoxDev *, int, void *, void *, const BoxState &, Queues *, int,
                            hipStream_t);
}  // namespace batch
}  // namespace evh

using namespace evh;
using namespace evh::batch;

struct evh_batch {
  int nbox = 0;
  std::vector<evh_config> cfgs;
  std::vector<KParams> kps;
  std::vector<long long> ncell, qoff;
  long long total = 0;
  int precision = EVH_FP64;
  size_t qelem = 8;
  int device = 0, num_cu = 256;
  hipStream_t stream = nullptr;
  void *qbuf[2] = {nullptr, nullptr};
  std::vector<int> side;  // host copy of bs.side, refreshed after every run
  void *snap = nullptr;  // autotuner snapshot of the state
  evh::Xfer xfer;  // host transfers (pinned, double-buffered chunks)
  BoxDev *boxes_d = nullptr;
  int nitems = 0;
  Queues *qs_d = nullptr;
  BoxState bs{};
  std::vector<char> cfl_needed;
  // launch shape
  int tile_ty = 12, wpe_hi = 1, zchunk = 64, blocks_per_cu = 2;
  // one step of the ensemble, captured once (the kernels read each box's
  // buffer side from device memory)
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string log;
  int err_code = 0, err_box = -1, err_cell[3] = {0, 0, 0};
  std::string err_msg;
};

namespace {

// what every box of a batch must share and support
int check_configs(const evh_config *cfgs, int nbox) {
  const evh_config &c0 = cfgs[0];
  for (int i = 0; i < nbox; i++) {
    const evh_config &c = cfgs[i];
    if (!fused_supported(c) || c.nz < 2) return EVH_ERR_CONFIG;  // 3-D boxes only
    if (c.conduction || c.gravity || c.hydro_off) return EVH_ERR_CONFIG;  // hydro step only
    if (c.solver != 1 && c.solver != 2) return EVH_ERR_SOLVER;
    if (c.limiter < 1 || c.limiter > 3) return EVH_ERR_LIMITER;
    if (c.solver != c0.solver || c.limiter != c0.limiter || c.precision != c0.precision ||
        c.secondorder != c0.secondorder || c.dualenergy != c0.dualenergy || c.boundtype < 1 || c.boundtype > 3 ||
        (c.precision != EVH_FP64 && c.precision != EVH_FP32_MIXED))
      return EVH_ERR_CONFIG;  // one kernel variant for the whole ensemble
  }
  return EVH_OK;
}

int bfail(evh_batch *b, int code, const char *msg, int box = -1) {
  b->err_code = code;
  b->err_box = box;
  b->err_msg = msg ? msg : "";
  return code;
}

#define BHIP(x)                                                                    \
  do {                                                                             \
    hipError_t e__ = (x);                                                          \
    if (e__ != hipSuccess) return bfail(b, EVH_ERR_HIP, hipGetErrorString(e__));   \
  } while (0)

template <typename S>
__global__ void k_b_aos_to_soa(const double *__restrict__ src, S *__restrict__ dst, long long cells, long long base,
                               long long n) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < cells; c += (long long)gridDim.x * 256)
#pragma unroll
    for (int m = 0; m < 6; m++) dst[m * n + base + c] = S(src[6 * c + m]);
}

template <typename S>
__global__ void k_b_soa_to_aos(const S *__restrict__ src, double *__restrict__ dst, long long cells, long long base,
                               long long n) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < cells; c += (long long)gridDim.x * 256)
#pragma unroll
    for (int m = 0; m < 6; m++) dst[6 * c + m] = double(src[m * n + base + c]);
}

__global__ void k_b_reset_status(Status *st, int box) {
  if (threadIdx.x == 0) {
    st[box].code = 0;
    st[box].cell = 0;
    st[box].next_code = 0;
    st[box].next_cell = 0;
    st[box].vmin_bits = (unsigned long long)__double_as_longlong(1e6);
  }
}

int gridsz(long long work) {
  long long g = (work + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// Descriptors of nbox boxes (box i at element offset 6 qoff[i]) for tile
// height ty and z chunk zc, and the per-XCD work queues over their tiles;
// returns the number of work items.
int describe_boxes(const evh_config *cfgs, const KParams *kps, const long long *ncell, const long long *qoff,
                   int nbox, int ty, int zc, BoxDev *out, Queues *q) {
  int items = 0;
  for (int i = 0; i < nbox; i++) {
    const evh_config &c = cfgs[i];
    Geo g;
    g.nx = c.nx;
    g.ny = c.ny;
    g.nz = c.nz;
    g.n = ncell[i];
    g.zchunk = zc > g.nz ? g.nz : zc;
    g.tiles_x = (g.nx + core::TX - 1) / core::TX;
    g.tiles_y = (g.ny + ty - 1) / ty;
    g.nchunks = (g.nz + g.zchunk - 1) / g.zchunk;
    g.ntiles = g.tiles_x * g.tiles_y * g.nchunks;
    g.zlo = g.zhi = nullptr;  // every box is a whole grid
    out[i].P = kps[i];
    out[i].G = g;
    out[i].qoff = 6 * qoff[i];
    out[i].tile0 = items;
    out[i].pad = 0;
    items += g.ntiles;
  }
  *q = Queues{};
  for (int j = 0; j <= kXcds; j++) q->start[j] = (int)((long long)items * j / kXcds);
  return items;
}

// descriptors and work queues for the current launch shape
int rebuild_items(evh_batch *b) {
  std::vector<BoxDev> h(b->nbox);
  Queues q;
  b->nitems = describe_boxes(b->cfgs.data(), b->kps.data(), b->ncell.data(), b->qoff.data(), b->nbox, b->tile_ty,
                             b->zchunk, h.data(), &q);
  BHIP(hipMemcpy(b->boxes_d, h.data(), sizeof(BoxDev) * b->nbox, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->qs_d, &q, sizeof(Queues), hipMemcpyHostToDevice));
  if (b->gexec) {
    hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  return EVH_OK;
}

hipError_t launch_step(evh_batch *b) {
  const int nb = b->nbox;
  hipLaunchKernelGGL(k_batch_dt, dim3((nb + 255) / 256), dim3(256), 0, b->stream, nb, b->bs, b->qs_d);
  const int blocks = b->blocks_per_cu * b->num_cu;
  const evh_config &c = b->cfgs[0];
  if (b->precision == EVH_FP32_MIXED)
    return launch_batch_f32(c.solver, c.limiter, b->tile_ty, b->wpe_hi != 0, b->boxes_d, nb, b->qbuf[0], b->qbuf[1],
                            b->bs, b->qs_d, blocks, b->stream);
  return launch_batch_f64(c.solver, c.limiter, b->tile_ty, b->wpe_hi != 0, b->boxes_d, nb, b->qbuf[0], b->qbuf[1],
                          b->bs, b->qs_d, blocks, b->stream);
}

int capture_graph(evh_batch *b) {
  if (b->gexec) return EVH_OK;
  hipGraph_t g = nullptr;
  BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = launch_step(b);
  hipError_t ce = hipStreamEndCapture(b->stream, &g);
  if (le != hipSuccess) return bfail(b, EVH_ERR_HIP, hipGetErrorString(le));
  BHIP(ce);
  hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  BHIP(ie);
  return EVH_OK;
}

// element pointer of box i's current state
void *box_q(evh_batch *b, int i) {
  char *base = (char *)b->qbuf[b->side[i]];
  return base + b->qelem * 6 * b->qoff[i];
}

int refresh_sides(evh_batch *b) {
  BHIP(hipMemcpy(b->side.data(), b->bs.side, 4 * b->nbox, hipMemcpyDeviceToHost));
  return EVH_OK;
}

// first-step CFL for boxes whose state was (re)uploaded
int prepare(evh_batch *b) {
  for (int i = 0; i < b->nbox; i++) {
    if (!b->cfl_needed[i]) continue;
    hipLaunchKernelGGL(k_b_reset_status, dim3(1), dim3(64), 0, b->stream, b->bs.st, i);
    const long long n = b->ncell[i];
    if (b->precision == EVH_FP32_MIXED)
      hipLaunchKernelGGL(k_batch_cfl<float>, dim3(gridsz(n)), dim3(256), 0, b->stream, b->kps[i], n,
                         (const float *)box_q(b, i), b->bs.st + i);
    else
      hipLaunchKernelGGL(k_batch_cfl<double>, dim3(gridsz(n)), dim3(256), 0, b->stream, b->kps[i], n,
                         (const double *)box_q(b, i), b->bs.st + i);
    BHIP(hipGetLastError());
    const int one = 1;
    BHIP(hipMemcpyAsync(b->bs.go + i, &one, sizeof(int), hipMemcpyHostToDevice, b->stream));
    b->cfl_needed[i] = 0;
  }
  BHIP(hipStreamSynchronize(b->stream));
  return EVH_OK;
}

// surface the first box error after a run
int check_boxes(evh_batch *b) {
  std::vector<Status> st(b->nbox);
  BHIP(hipMemcpy(st.data(), b->bs.st, sizeof(Status) * b->nbox, hipMemcpyDeviceToHost));
  for (int i = 0; i < b->nbox; i++) {
    const Status &s = st[i];
    const int code = s.code ? s.code : (int)s.next_code;
    if (!code) continue;
    const unsigned long long cell = (s.code ? s.cell : s.next_cell) & ((1ULL << 40) - 1);
    const long long nx = b->cfgs[i].nx, ny = b->cfgs[i].ny;
    b->err_cell[0] = (int)(cell % nx) + 1;
    b->err_cell[1] = (int)((cell / nx) % ny) + 1;
    b->err_cell[2] = (int)(cell / (nx * ny)) + 1;
    char msg[128];
    std::snprintf(msg, sizeof msg, "box %d stopped with status %d", i, code);
    return bfail(b, code, msg, i);
  }
  return EVH_OK;
}

struct Snapshot {
  std::vector<double> t, tend, dt;
  std::vector<int> go, steps, err_step, side;
  std::vector<Status> st;
};

// every box's current state into (or out of) the snapshot buffer, at the
// box's own offset
int copy_boxes(evh_batch *b, bool save) {
  for (int i = 0; i < b->nbox; i++) {
    char *snap = (char *)b->snap + b->qelem * 6 * b->qoff[i];
    void *cur = box_q(b, i);
    const size_t bytes = b->qelem * 6 * b->ncell[i];
    BHIP(hipMemcpyAsync(save ? (void *)snap : cur, save ? cur : (void *)snap, bytes, hipMemcpyDeviceToDevice,
                        b->stream));
  }
  return EVH_OK;
}

int save_state(evh_batch *b, Snapshot &s) {
  const int nb = b->nbox;
  if (!b->snap) BHIP(hipMalloc(&b->snap, b->qelem * 6 * b->total));
  int r = copy_boxes(b, true);
  if (r) return r;
  s.t.resize(nb); s.tend.resize(nb); s.dt.resize(nb); s.go.resize(nb); s.steps.resize(nb); s.err_step.resize(nb);
  s.st.resize(nb);
  BHIP(hipMemcpy(s.t.data(), b->bs.t, 8 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.tend.data(), b->bs.tend, 8 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.dt.data(), b->bs.dt, 8 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.go.data(), b->bs.go, 4 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.steps.data(), b->bs.steps, 4 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.err_step.data(), b->bs.err_step, 4 * nb, hipMemcpyDeviceToHost));
  BHIP(hipMemcpy(s.st.data(), b->bs.st, sizeof(Status) * nb, hipMemcpyDeviceToHost));
  s.side = b->side;
  return EVH_OK;
}

int restore_state(evh_batch *b, const Snapshot &s) {
  const int nb = b->nbox;
  b->side = s.side;
  BHIP(hipMemcpy(b->bs.side, s.side.data(), 4 * nb, hipMemcpyHostToDevice));
  int r = copy_boxes(b, false);
  if (r) return r;
  BHIP(hipMemcpy(b->bs.t, s.t.data(), 8 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.tend, s.tend.data(), 8 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.dt, s.dt.data(), 8 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.go, s.go.data(), 4 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.steps, s.steps.data(), 4 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.err_step, s.err_step.data(), 4 * nb, hipMemcpyHostToDevice));
  BHIP(hipMemcpy(b->bs.st, s.st.data(), sizeof(Status) * nb, hipMemcpyHostToDevice));
  BHIP(hipStreamSynchronize(b->stream));
  return EVH_OK;
}

}  // namespace

extern "C" {

int evh_batch_create(const evh_config *cfgs, int nbox, evh_batch **out) {
  if (!cfgs || !out || nbox < 1) return EVH_ERR_CONFIG;
  *out = nullptr;
  const int v = check_configs(cfgs, nbox);
  if (v) return v;
  const evh_config &c0 = cfgs[0];
  evh_batch *b = new (std::nothrow) evh_batch();
  if (!b) return EVH_ERR_HIP;
  b->nbox = nbox;
  b->cfgs.assign(cfgs, cfgs + nbox);
  b->precision = c0.precision;
  b->qelem = b->precision == EVH_FP32_MIXED ? 4 : 8;
  // default launch shape (the autotuner may pick another): 16x12 tiles, the
  // higher waves-per-SIMD budget, as many resident blocks per CU as that
  // occupancy allows (FP32: 4 waves/SIMD, FP64: 2)
  b->wpe_hi = 1;
  b->blocks_per_cu = b->precision == EVH_FP32_MIXED ? 4 : 2;
  long long off = 0;
  for (int i = 0; i < nbox; i++) {
    b->kps.push_back(make_kparams(cfgs[i]));
    const long long n = (long long)cfgs[i].nx * cfgs[i].ny * cfgs[i].nz;
    b->ncell.push_back(n);
    b->qoff.push_back(off);
    off += n;
  }
  b->total = off;
  b->cfl_needed.assign(nbox, 0);
  auto bail = [&](int code) {
    evh_batch_destroy(b);
    return code;
  };
  hipDeviceProp_t prop;
  if (hipGetDevice(&b->device) != hipSuccess || hipGetDeviceProperties(&prop, b->device) != hipSuccess)
    return bail(EVH_ERR_HIP);
  b->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(EVH_ERR_HIP);
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return bail(EVH_ERR_HIP);
  for (int i = 0; i < 2; i++)
    if (hipMalloc(&b->qbuf[i], b->qelem * 6 * b->total) != hipSuccess) return bail(EVH_ERR_HIP);
  long long maxn = 0;
  for (auto n : b->ncell) maxn = n > maxn ? n : maxn;
  if (b->xfer.init(maxn) != hipSuccess) return bail(EVH_ERR_HIP);
  if (hipMalloc(&b->boxes_d, sizeof(BoxDev) * nbox) != hipSuccess) return bail(EVH_ERR_HIP);
  if (hipMalloc(&b->qs_d, sizeof(Queues)) != hipSuccess) return bail(EVH_ERR_HIP);
  BoxState &s = b->bs;
  if (hipMalloc(&s.t, 8 * nbox) || hipMalloc(&s.tend, 8 * nbox) || hipMalloc(&s.dt, 8 * nbox) ||
      hipMalloc(&s.cfl, 8 * nbox) || hipMalloc(&s.go, 4 * nbox) || hipMalloc(&s.active, 4 * nbox) ||
      hipMalloc(&s.steps, 4 * nbox) || hipMalloc(&s.err_step, 4 * nbox) || hipMalloc(&s.side, 4 * nbox) ||
      hipMalloc(&s.st, sizeof(Status) * nbox))
    return bail(EVH_ERR_HIP);
  b->side.assign(nbox, 0);
  std::vector<double> zeros(nbox, 0.0), big(nbox, 1e99), cfl(nbox);
  std::vector<int> izero(nbox, 0), ineg(nbox, -1);
  for (int i = 0; i < nbox; i++) cfl[i] = cfgs[i].cfl;
  std::vector<Status> st(nbox);
  for (auto &x : st) {
    std::memset(&x, 0, sizeof x);
    x.vmin_bits = 0;
  }
  if (hipMemcpy(s.t, zeros.data(), 8 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.tend, big.data(), 8 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.dt, zeros.data(), 8 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.cfl, cfl.data(), 8 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.go, izero.data(), 4 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.active, izero.data(), 4 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.steps, izero.data(), 4 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.err_step, ineg.data(), 4 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.side, izero.data(), 4 * nbox, hipMemcpyHostToDevice) ||
      hipMemcpy(s.st, st.data(), sizeof(Status) * nbox, hipMemcpyHostToDevice))
    return bail(EVH_ERR_HIP);
  if (rebuild_items(b)) return bail(EVH_ERR_HIP);
  *out = b;
  return EVH_OK;
}

int evh_batch_destroy(evh_batch *b) {
  if (!b) return EVH_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  for (auto p : b->qbuf)
    if (p) hipFree(p);
  if (b->snap) hipFree(b->snap);
  b->xfer.release();
  if (b->boxes_d) hipFree(b->boxes_d);
  if (b->qs_d) hipFree(b->qs_d);
  void *arrs[] = {b->bs.t,     b->bs.tend,  b->bs.dt,       b->bs.cfl,  b->bs.go,
                  b->bs.active, b->bs.steps, b->bs.err_step, b->bs.side, b->bs.st};
  for (auto p : arrs)
    if (p) hipFree(p);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return EVH_OK;
}

int evh_batch_upload_q(evh_batch *b, int box, const double *q) {
  if (!b || !q || box < 0 || box >= b->nbox) return EVH_ERR_STATE;
  const long long n = b->ncell[box];
  const bool f32 = b->precision == EVH_FP32_MIXED;
  void *dst = box_q(b, box);
  hipStream_t s = b->stream;
  BHIP(b->xfer.upload(q, n, s, [=](const double *stage, long long cells, long long base) {
    if (f32)
      hipLaunchKernelGGL(k_b_aos_to_soa<float>, dim3(gridsz(cells)), dim3(256), 0, s, stage, (float *)dst, cells, base,
                         n);
    else
      hipLaunchKernelGGL(k_b_aos_to_soa<double>, dim3(gridsz(cells)), dim3(256), 0, s, stage, (double *)dst, cells,
                         base, n);
    return hipGetLastError();
  }));
  b->cfl_needed[box] = 1;
  return EVH_OK;
}

int evh_batch_download_q(evh_batch *b, int box, double *q) {
  if (!b || !q || box < 0 || box >= b->nbox) return EVH_ERR_STATE;
  const long long n = b->ncell[box];
  const bool f32 = b->precision == EVH_FP32_MIXED;
  const void *src = box_q(b, box);
  hipStream_t s = b->stream;
  BHIP(b->xfer.download(q, n, s, [=](double *stage, long long cells, long long base) {
    if (f32)
      hipLaunchKernelGGL(k_b_soa_to_aos<float>, dim3(gridsz(cells)), dim3(256), 0, s, (const float *)src, stage, cells,
                         base, n);
    else
      hipLaunchKernelGGL(k_b_soa_to_aos<double>, dim3(gridsz(cells)), dim3(256), 0, s, (const double *)src, stage,
                         cells, base, n);
    return hipGetLastError();
  }));
  return EVH_OK;
}

int evh_batch_set_time(evh_batch *b, const double *t, const double *tend) {
  if (!b) return EVH_ERR_STATE;
  if (t) BHIP(hipMemcpy(b->bs.t, t, 8 * b->nbox, hipMemcpyHostToDevice));
  if (tend) BHIP(hipMemcpy(b->bs.tend, tend, 8 * b->nbox, hipMemcpyHostToDevice));
  return EVH_OK;
}

int evh_batch_run(evh_batch *b, int nsteps, int use_graph, float *ms) {
  if (!b || nsteps < 0) return EVH_ERR_STATE;
  int r = prepare(b);
  if (r) return r;
  if (use_graph && (r = capture_graph(b))) return r;
  BHIP(hipEventRecord(b->ev0, b->stream));
  for (int i = 0; i < nsteps; i++) {
    if (use_graph) BHIP(hipGraphLaunch(b->gexec, b->stream));
    else BHIP(launch_step(b));
  }
  BHIP(hipEventRecord(b->ev1, b->stream));
  BHIP(hipEventSynchronize(b->ev1));
  if (ms) BHIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  if ((r = refresh_sides(b))) return r;
  return check_boxes(b);
}

int evh_batch_times(evh_batch *b, double *t, double *dt, int *steps, int *go) {
  if (!b) return EVH_ERR_STATE;
  BHIP(hipStreamSynchronize(b->stream));
  if (t) BHIP(hipMemcpy(t, b->bs.t, 8 * b->nbox, hipMemcpyDeviceToHost));
  if (dt) BHIP(hipMemcpy(dt, b->bs.dt, 8 * b->nbox, hipMemcpyDeviceToHost));
  if (steps) BHIP(hipMemcpy(steps, b->bs.steps, 4 * b->nbox, hipMemcpyDeviceToHost));
  if (go) BHIP(hipMemcpy(go, b->bs.go, 4 * b->nbox, hipMemcpyDeviceToHost));
  return EVH_OK;
}

int evh_batch_set_tuning(evh_batch *b, int tile_ty, int wpe_hi, int zchunk, int blocks_per_cu) {
  if (!b) return EVH_ERR_STATE;
  if ((tile_ty != 8 && tile_ty != 12) || (wpe_hi != 0 && wpe_hi != 1) || zchunk < 1 || blocks_per_cu < 1 ||
      blocks_per_cu > 8)
    return bfail(b, EVH_ERR_CONFIG, "tile_ty 8|12, wpe_hi 0|1, zchunk >= 1, blocks_per_cu 1..8");
  b->tile_ty = tile_ty;
  b->wpe_hi = wpe_hi;
  b->zchunk = zchunk;
  b->blocks_per_cu = blocks_per_cu;
  return rebuild_items(b);
}

int evh_batch_get_tuning(evh_batch *b, int *tile_ty, int *wpe_hi, int *zchunk, int *blocks_per_cu) {
  if (!b || !tile_ty || !wpe_hi || !zchunk || !blocks_per_cu) return EVH_ERR_STATE;
  *tile_ty = b->tile_ty;
  *wpe_hi = b->wpe_hi;
  *zchunk = b->zchunk;
  *blocks_per_cu = b->blocks_per_cu;
  return EVH_OK;
}

int evh_batch_autotune(evh_batch *b, int trial_steps, char *log, int loglen) {
  if (!b || trial_steps < 1) return EVH_ERR_STATE;
  int r = prepare(b);
  if (r) return r;
  Snapshot snap;
  if ((r = save_state(b, snap))) return r;
  const int tys[] = {8, 12}, zcs[] = {32, 64, 128}, bpcs[] = {1, 2, 3, 4};
  double best = 1e300;
  int bt[4] = {b->tile_ty, b->wpe_hi, b->zchunk, b->blocks_per_cu};
  std::string text = "# batch autotune: nbox=" + std::to_string(b->nbox) + " cells=" + std::to_string(b->total) +
                     " trial_steps=" + std::to_string(trial_steps) + "\n";
  for (int ty : tys)
    for (int hi = 0; hi <= 1; hi++)
      for (int zc : zcs)
        for (int bpc : bpcs) {
          if ((r = evh_batch_set_tuning(b, ty, hi, zc, bpc))) return r;
          if ((r = restore_state(b, snap))) return r;
          float ms = 0.f;
          if ((r = evh_batch_run(b, 1, 0, &ms))) return r;  // warm-up step
          if ((r = evh_batch_run(b, trial_steps, 0, &ms))) return r;
          const double per = ms / trial_steps;
          char line[160];
          std::snprintf(line, sizeof line, "tile_ty=%d wpe_hi=%d zchunk=%d blocks_per_cu=%d ms_per_step=%.4f\n", ty, hi,
                        zc, bpc, per);
          text += line;
          if (per < best) {
            best = per;
            bt[0] = ty; bt[1] = hi; bt[2] = zc; bt[3] = bpc;
          }
        }
  char line[160];
  std::snprintf(line, sizeof line, "best: tile_ty=%d wpe_hi=%d zchunk=%d blocks_per_cu=%d ms_per_step=%.4f\n", bt[0],
                bt[1], bt[2], bt[3], best);
  text += line;
  b->log = text;
  if ((r = evh_batch_set_tuning(b, bt[0], bt[1], bt[2], bt[3]))) return r;
  if ((r = restore_state(b, snap))) return r;
  if (log && loglen > 0) {
    std::strncpy(log, text.c_str(), (size_t)loglen - 1);
    log[loglen - 1] = 0;
  }
  return EVH_OK;
}

int evh_batch_last_error(evh_batch *b, int *code, int *box, int cell[3], char *msg, int msglen) {
  if (!b) return EVH_ERR_STATE;
  if (code) *code = b->err_code;
  if (box) *box = b->err_box;
  if (cell)
    for (int i = 0; i < 3; i++) cell[i] = b->err_cell[i];
  if (msg && msglen > 0) {
    std::strncpy(msg, b->err_msg.c_str(), (size_t)msglen - 1);
    msg[msglen - 1] = 0;
  }
  return EVH_OK;
}


// ---------------------------------------------------------------- allocation-free form
// evh_batch_run_noalloc: the batched step for callers that cannot allocate
// device memory.  Everything the kernels need beyond the caller's two state
// buffers -- box descriptors, work queues, per-box time, dt, flags and status
// blocks -- lives in `scratch`, a fixed device array of this file (one copy
// per GPU, for the life of the process).  Every call rewrites all of it
// before the first kernel reads it; calls that overlap on one GPU race on it.
constexpr int kNoAllocMaxBoxes = 64;

struct NoAllocScratch {
  BoxDev boxes[kNoAllocMaxBoxes];
  Queues qs;
  double t[kNoAllocMaxBoxes], tend[kNoAllocMaxBoxes], dt[kNoAllocMaxBoxes], cfl[kNoAllocMaxBoxes];
  int go[kNoAllocMaxBoxes], active[kNoAllocMaxBoxes], steps[kNoAllocMaxBoxes], err_step[kNoAllocMaxBoxes];
  int side[kNoAllocMaxBoxes];
  Status st[kNoAllocMaxBoxes];
};
constexpr size_t kScratchBytes = 48 * 1024;
static_assert(sizeof(NoAllocScratch) <= kScratchBytes, "scratch too small for its layout");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *scratch_ptr[kMaxDevices] = {};

// the device address of this GPU's copy of `scratch`, looked up once per
// device (the host-side name is not a device address)
hipError_t scratch_address(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return hipSuccess;
}

int na_fail(char *err, int errlen, int code, const char *msg) {
  if (err && errlen > 0) {
    std::strncpy(err, msg, (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
  return code;
}
}  // namespace

int evh_batch_run_noalloc(const evh_config *cfgs, int nbox, void *q0, void *q1, int *side, double *t,
                          const double *tend, int *steps, int *go, int nsteps, char *err, int errlen) {
  if (err && errlen > 0) err[0] = 0;
  if (!cfgs || !q0 || !q1 || !side || !t || !tend || !steps || !go || nsteps < 0)
    return na_fail(err, errlen, EVH_ERR_STATE, "null argument or negative nsteps");
  if (nbox < 1 || nbox > kNoAllocMaxBoxes) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "nbox %d outside 1..%d, the boxes the allocation-free scratch (%zu bytes) holds",
                  nbox, kNoAllocMaxBoxes, kScratchBytes);
    return na_fail(err, errlen, EVH_ERR_CONFIG, msg);
  }
  const int v = check_configs(cfgs, nbox);
  if (v) return na_fail(err, errlen, v, "unsupported box configuration");
  for (int i = 0; i < nbox; i++)
    if (side[i] != 0 && side[i] != 1) return na_fail(err, errlen, EVH_ERR_STATE, "side must be 0 or 1");
  void *sp = nullptr;
  hipError_t e = scratch_address(&sp);
  if (e != hipSuccess) return na_fail(err, errlen, EVH_ERR_HIP, hipGetErrorString(e));
  NoAllocScratch *dev = (NoAllocScratch *)sp;

  // the whole scratch image, formed on the host stack (no heap either)
  NoAllocScratch img;
  std::memset(&img, 0, sizeof img);
  KParams kps[kNoAllocMaxBoxes];
  long long ncell[kNoAllocMaxBoxes], qoff[kNoAllocMaxBoxes];
  long long off = 0;
  const bool f32 = cfgs[0].precision == EVH_FP32_MIXED;
  for (int i = 0; i < nbox; i++) {
    kps[i] = make_kparams(cfgs[i]);
    ncell[i] = (long long)cfgs[i].nx * cfgs[i].ny * cfgs[i].nz;
    qoff[i] = off;
    off += ncell[i];
  }
  // the launch shape evh_batch_create starts with
  const int ty = 12, zc = 64, wpe_hi = 1, bpc = f32 ? 4 : 2;
  describe_boxes(cfgs, kps, ncell, qoff, nbox, ty, zc, img.boxes, &img.qs);
  for (int i = 0; i < nbox; i++) {
    img.t[i] = t[i];
    img.tend[i] = tend[i];
    img.dt[i] = 0.0;
    img.cfl[i] = cfgs[i].cfl;
    img.go[i] = go[i] ? 1 : 0;
    img.steps[i] = steps[i];
    img.err_step[i] = -1;
    img.side[i] = side[i];
  }
  BoxState bs;
  bs.t = dev->t;
  bs.tend = dev->tend;
  bs.dt = dev->dt;
  bs.cfl = dev->cfl;
  bs.go = dev->go;
  bs.active = dev->active;
  bs.steps = dev->steps;
  bs.err_step = dev->err_step;
  bs.side = dev->side;
  bs.st = dev->st;
  int devid = 0, cu = 256;
  if ((e = hipGetDevice(&devid)) == hipSuccess)
    e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, devid);
  if (e == hipSuccess) e = hipMemcpy(dev, &img, sizeof img, hipMemcpyHostToDevice);
  // the first CFL limit of every running box from its current state
  for (int i = 0; i < nbox && e == hipSuccess; i++) {
    if (!img.go[i]) continue;
    hipLaunchKernelGGL(k_b_reset_status, dim3(1), dim3(64), 0, 0, bs.st, i);
    const long long n = ncell[i];
    if (f32)
      hipLaunchKernelGGL(k_batch_cfl<float>, dim3(gridsz(n)), dim3(256), 0, 0, kps[i], n,
                         (const float *)(side[i] ? q1 : q0) + 6 * qoff[i], bs.st + i);
    else
      hipLaunchKernelGGL(k_batch_cfl<double>, dim3(gridsz(n)), dim3(256), 0, 0, kps[i], n,
                         (const double *)(side[i] ? q1 : q0) + 6 * qoff[i], bs.st + i);
    e = hipGetLastError();
  }
  const evh_config &c0 = cfgs[0];
  for (int k = 0; k < nsteps && e == hipSuccess; k++) {
    hipLaunchKernelGGL(k_batch_dt, dim3((nbox + 255) / 256), dim3(256), 0, 0, nbox, bs, &dev->qs);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    if (f32)
      e = launch_batch_f32(c0.solver, c0.limiter, ty, wpe_hi != 0, dev->boxes, nbox, q0, q1, bs, &dev->qs,
                           bpc * cu, 0);
    else
      e = launch_batch_f64(c0.solver, c0.limiter, ty, wpe_hi != 0, dev->boxes, nbox, q0, q1, bs, &dev->qs,
                           bpc * cu, 0);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(&img, dev, sizeof img, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return na_fail(err, errlen, EVH_ERR_HIP, hipGetErrorString(e));
  int code = EVH_OK;
  for (int i = 0; i < nbox; i++) {
    t[i] = img.t[i];
    steps[i] = img.steps[i];
    go[i] = img.go[i];
    side[i] = img.side[i];
    const Status &st = img.st[i];
    const int c = st.code ? st.code : (int)st.next_code;
    if (c && !code) {
      code = c;
      char msg[96];
      std::snprintf(msg, sizeof msg, "box %d stopped with status %d", i, c);
      na_fail(err, errlen, c, msg);
    }
  }
  return code;
}

}  // extern "C"
#endif
