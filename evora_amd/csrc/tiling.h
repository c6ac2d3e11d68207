// tiling.h -- the block-to-tile mapping shared by the tile-marching kernels
// (fused sweep, conduction, gravity force) and the XCD count the batched
// work queues are split by.
#pragma once

#include <hip/hip_runtime.h>

namespace evh {

// Blocks are dispatched round-robin over the 8 XCDs of the chip, each with
// its own L2.  Block b runs on XCD b % 8; giving XCD j the contiguous run of
// tiles [j*T/8, (j+1)*T/8) (x fastest, then y, then z chunk) keeps the halo
// columns a block shares with its x and y neighbours in one L2.  Returns -1
// for the padding blocks of a grid rounded up to a multiple of 8.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int j = b % kXcds, i = b / kXcds;
  const int base = ntiles / kXcds, rem = ntiles % kXcds;
  const int count = base + (j < rem ? 1 : 0);
  if (i >= count) return -1;
  return j * base + (j < rem ? j : rem) + i;
}

// blocks to launch for ntiles tiles (a multiple of the XCD count)
inline int xcd_blocks(int ntiles) { return kXcds * ((ntiles + kXcds - 1) / kXcds); }

// Chunk lengths when the tuning leaves them open.  A block's march is
// serial, so on a small grid long chunks leave most of the GPU idle and a
// pass takes the march's latency; short chunks cost a few planes more of
// loads and primitives per chunk.  The results do not depend on the chunk.
//
// z_chunk: planes per block of a 3-D march (the fused step, conduction, the
// gravity force) with `columns` column tiles per plane: `start`, halved
// while the grid would give fewer than 2048 blocks, but not below nz/8 (a
// power of two, 4 at least).  Measured fastest on MI355X: the fused step of
// the Kelvin-Helmholtz cube at 8 planes for 96^3, 16 for 128^3 and 192^3,
// 32 for 256^3, 64 for 512^3 and 32 for a 512 x 512 x 64 slab; the
// conduction pass of a 128^3 hot blob 0.082 ms at 16 planes against 0.207 ms
// at 4 (1024 blocks of 4 planes).
inline int z_chunk(long long columns, int nz, int start) {
  int lo = 4;
  while (lo < start && 2 * lo * 8 <= nz) lo *= 2;
  int c = start;
  while (c > lo && columns * ((nz + c - 1) / c) < 2048) c /= 2;
  return c;
}

// y_chunk: rows per block of the 2-D march with `strips` row strips:
// `start`, halved down to 4 while the grid would give fewer than 1024
// blocks, and to 2 while it would give fewer than 128 (a row is a block's
// whole strip, so short chunks cost little: the step of para/kevinhelmholtz's
// 128^2 grid measured 0.0315 ms at 2 rows against 0.0361 at 4; 512^2 is
// fastest at 4 rows, 8192^2 at 32).
inline int y_chunk(long long strips, int ny, int start) {
  int c = start;
  while (c > 4 && strips * ((ny + c - 1) / c) < 1024) c /= 2;
  while (c > 2 && strips * ((ny + c - 1) / c) < 128) c /= 2;
  return c;
}

}  // namespace evh
