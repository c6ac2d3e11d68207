// transfer.h -- host <-> device copies of q in the reference layout.
//
// q crosses PCIe in chunks of at most kChunkBytes through two pinned host
// buffers and two device staging buffers.  Per chunk the device runs a
// layout kernel (AoS <-> SoA, given by the caller) and one DMA; the host
// copies between the caller's (pageable) array and the pinned buffer on a few
// threads.  The host copy of one chunk overlaps the DMA and the kernel of the
// next, so a transfer runs at the slower of PCIe and host memcpy instead of
// their sum plus a synchronisation per chunk.  The pinned buffers are
// allocated on first use (hipHostMalloc).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace evh {

constexpr long long kChunkBytes = 64LL << 20;  // per pinned buffer
constexpr long long kCellBytes = 6 * 8;       // one cell of q in the reference layout (6 doubles)

// memcpy on up to 8 threads for large blocks
inline void par_copy(void *dst, const void *src, size_t bytes) {
  const size_t kMin = 8u << 20;
  unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  if (bytes < 2 * kMin || nt == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  nt = (unsigned)std::min<size_t>(nt, bytes / kMin);
  const size_t part = (bytes / nt + 63) & ~size_t(63);
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; t++) {
    const size_t off = t * part;
    if (off >= bytes) break;
    const size_t len = std::min(part, bytes - off);
    pool.emplace_back([=] { std::memcpy((char *)dst + off, (const char *)src + off, len); });
  }
  std::memcpy(dst, src, std::min(part, bytes));
  for (auto &th : pool) th.join();
}

struct Xfer {
  double *dstage = nullptr;              // device: 2 * chunk_cells * 6 doubles
  double *pin[2] = {nullptr, nullptr};   // pinned host: chunk_cells * 6 doubles each
  hipEvent_t ev[2] = {nullptr, nullptr};
  long long chunk_cells = 0;

  // device staging for grids of up to max_cells cells
  hipError_t init(long long max_cells) {
    chunk_cells = std::max(1LL, std::min(max_cells, kChunkBytes / kCellBytes));
    hipError_t e = hipMalloc(&dstage, 2 * chunk_cells * kCellBytes);
    for (auto &x : ev)
      if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
    return e;
  }
  long long device_bytes() const { return 2 * chunk_cells * kCellBytes; }
  void release() {
    if (dstage) hipFree(dstage);
    for (auto p : pin)
      if (p) hipHostFree(p);
    for (auto x : ev)
      if (x) hipEventDestroy(x);
    dstage = nullptr;
    pin[0] = pin[1] = nullptr;
    ev[0] = ev[1] = nullptr;
  }
  hipError_t ensure_pinned() {
    for (auto &p : pin)
      if (!p) {
        hipError_t e = hipHostMalloc((void **)&p, chunk_cells * kCellBytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
      }
    return hipSuccess;
  }
  double *half(int b) const { return dstage + (long long)b * 6 * chunk_cells; }

  // host q (n cells) -> device; launch(staging, cells, base) enqueues the
  // layout kernel of one chunk on stream s
  template <typename F>
  hipError_t upload(const double *q, long long n, hipStream_t s, F launch) {
    hipError_t e = ensure_pinned();
    if (e != hipSuccess) return e;
    long long i = 0;
    for (long long base = 0; base < n; base += chunk_cells, i++) {
      const int b = (int)(i & 1);
      const long long cells = std::min(chunk_cells, n - base);
      if (i >= 2 && (e = hipEventSynchronize(ev[b])) != hipSuccess) return e;  // pin[b] consumed
      par_copy(pin[b], q + 6 * base, cells * kCellBytes);
      if ((e = hipMemcpyAsync(half(b), pin[b], cells * kCellBytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
      if ((e = hipEventRecord(ev[b], s)) != hipSuccess) return e;
      if ((e = launch(half(b), cells, base)) != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
  }

  // device -> host q (n cells); launch as for upload
  template <typename F>
  hipError_t download(double *q, long long n, hipStream_t s, F launch) {
    hipError_t e = ensure_pinned();
    if (e != hipSuccess) return e;
    long long prev_base = -1, prev_cells = 0, i = 0;
    int prev_b = 0;
    for (long long base = 0; base < n; base += chunk_cells, i++) {
      const int b = (int)(i & 1);
      const long long cells = std::min(chunk_cells, n - base);
      // pin[b] held chunk i-2, copied out on the host in the previous iteration
      if ((e = launch(half(b), cells, base)) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(pin[b], half(b), cells * kCellBytes, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
      if ((e = hipEventRecord(ev[b], s)) != hipSuccess) return e;
      if (prev_base >= 0) {
        if ((e = hipEventSynchronize(ev[prev_b])) != hipSuccess) return e;
        par_copy(q + 6 * prev_base, pin[prev_b], prev_cells * kCellBytes);
      }
      prev_base = base;
      prev_cells = cells;
      prev_b = b;
    }
    if (prev_base >= 0) {
      if ((e = hipEventSynchronize(ev[prev_b])) != hipSuccess) return e;
      par_copy(q + 6 * prev_base, pin[prev_b], prev_cells * kCellBytes);
    }
    return hipSuccess;
  }
};

}  // namespace evh
