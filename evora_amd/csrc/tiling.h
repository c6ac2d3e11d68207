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

// Planes (rows) per block of a march along n with `columns` column tiles per
// plane: `start`, halved down to 4 while the grid would give fewer than 1024
// blocks.  A block's march is serial, so on a small grid long chunks leave
// most of the GPU idle and the pass takes the march's latency; short chunks
// cost a few planes more of loads and primitives per chunk.  The results do
// not depend on the chunk.
inline int auto_chunk(long long columns, int n, int start) {
  int c = start;
  while (c > 4 && columns * ((n + c - 1) / c) < 1024) c /= 2;
  return c;
}

}  // namespace evh
