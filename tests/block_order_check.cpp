// CPU check of the fused pass's block order (xadvcool_amd/csrc/lattice.h,
// launch_block / xcd_grid), for every order (linear, reversed, XCD-contiguous,
// both), slots S = 1..6 and blocks per slot bps = 1..600 plus the sizes of
// 148x128x32 lattices at 64, 128 and 256 cells per block:
//   * the launch blocks that are not padding visit every lattice block
//     s * bps + lb exactly once;
//   * linear order: launch block b works on lattice block b (reversed:
//     grid - 1 - b) and the grid has no padding;
//   * XCD order: launch block b (on XCD b % 8) works on the (b % 8)-th eighth
//     of its slot in forward passes and the (7 - b % 8)-th in reversed ones,
//     so every slot is spread over all eight XCDs in equal shares.
// Built with g++ by tests/test_block_order.py; prints "ok <cases>" or the
// first mismatch and exits non-zero.
#define __host__
#define __device__
#include <cstdio>
#include <vector>

#include "../xadvcool_amd/csrc/lattice.h"

using namespace xadv;

static int check(int bps, int S, int order) {
  const int G = bps * S;
  const int grid = (order & 1) ? xcd_grid(bps, S) : G;
  if (grid < G || grid % ((order & 1) ? 8 : 1)) {
    printf("grid %d for bps %d S %d order %d\n", grid, bps, S, order);
    return 1;
  }
  const int per = (bps + 7) / 8;
  std::vector<int> seen(G, 0);
  int pad = 0;
  for (int b = 0; b < grid; ++b) {
    const int m = launch_block(b, grid, bps, order);
    if (m >= G) {
      ++pad;
      continue;
    }
    if (m < 0 || seen[m]++) {
      printf("block %d -> %d twice or negative (bps %d S %d order %d)\n", b, m, bps, S, order);
      return 1;
    }
    if (!(order & 1)) {
      const int want = (order & 2) ? grid - 1 - b : b;
      if (m != want) {
        printf("linear order: block %d -> %d, want %d (bps %d S %d order %d)\n", b, m, want, bps, S, order);
        return 1;
      }
    } else {
      const int lb = m % bps, eighth = lb / per;
      const int xcd = b % 8, want = (order & 2) ? 7 - xcd : xcd;
      if (eighth != want) {
        printf("xcd order: block %d (XCD %d) -> slot block %d in eighth %d, want %d (bps %d S %d order %d)\n", b,
               xcd, lb, eighth, want, bps, S, order);
        return 1;
      }
    }
  }
  if (pad != grid - G) {
    printf("padding %d, want %d (bps %d S %d order %d)\n", pad, grid - G, bps, S, order);
    return 1;
  }
  return 0;
}

int main() {
  std::vector<int> sizes;
  for (int b = 1; b <= 600; ++b) sizes.push_back(b);
  const int cells = 148 * 128 * 32;
  for (int cpb : {64, 128, 256, 512}) sizes.push_back((cells + cpb - 1) / cpb);
  long cases = 0;
  for (int bps : sizes)
    for (int S = 1; S <= 6; ++S)
      for (int order = 0; order < 4; ++order) {
        if (check(bps, S, order)) return 1;
        ++cases;
      }
  for (int order = 0; order < 4; ++order) {  // the 64-design batch of 148x128x32 lattices
    if (check((cells + 255) / 256, 64, order)) return 1;
    ++cases;
  }
  printf("ok %ld\n", cases);
  return 0;
}
