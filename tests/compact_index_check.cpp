// CPU check of the compact fluid numbering used by the HIP pass
// (xadvcool_amd/csrc/lattice.h): for random masks, every non-solid cell's
// number equals its rank among non-solid cells in (x, y, z) order, and every
// neighbour number the pass uses (neighbour inside the lattice and
// non-solid) equals that neighbour's rank.  Also checks that the column
// indices read around every cell stay inside the column table.
// Built with g++ by tests/test_compact_index.py; prints "ok <cases>" or the
// first mismatch and exits non-zero.
#define __host__
#define __device__
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../xadvcool_amd/csrc/lattice.h"

using namespace xadv;

int main() {
  std::mt19937 rng(12345);
  const int shapes[][3] = {{3, 3, 1}, {4, 5, 2}, {7, 3, 32}, {9, 8, 31}, {12, 11, 16}, {5, 17, 32}, {3, 1, 5}};
  int cases = 0;
  for (const auto& sh : shapes)
    for (double solid : {0.0, 0.3, 0.7, 0.95, 1.0}) {
      const int nx = sh[0], ny = sh[1], nz = sh[2];
      const int64_t n = (int64_t)nx * ny * nz;
      std::vector<int8_t> mask(n);
      std::uniform_real_distribution<double> U(0.0, 1.0);
      for (auto& m : mask) m = U(rng) < solid ? BC_SOLID : (int8_t)(U(rng) < 0.5 ? BC_FLUID : BC_OUTLET);
      std::vector<ColInfo> cols((size_t)nx * ny);
      const int64_t nf = build_columns(mask.data(), nx, ny, nz, cols.data());
      std::vector<int64_t> rank(n, -1);
      int64_t r = 0;
      for (int64_t c = 0; c < n; ++c)
        if (mask[c] != BC_SOLID) rank[c] = r++;
      if (r != nf) {
        printf("count mismatch %lld vs %lld\n", (long long)r, (long long)nf);
        return 1;
      }
      const int ncol = nx * ny;
      for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
          for (int z = 0; z < nz; ++z) {
            const int64_t c = ((int64_t)x * ny + y) * nz + z;
            if (mask[c] == BC_SOLID) continue;
            const int q = x * ny + y;
            for (int dx = -1; dx <= 1; ++dx)
              for (int dy = -1; dy <= 1; ++dy) {
                int t = q + dx * ny + dy;
                t = t < 0 ? 0 : (t >= ncol ? ncol - 1 : t);
                if (t < 0 || t >= ncol) {
                  printf("column index out of range\n");
                  return 1;
                }
              }
            ColInfo col[3][3];
            around_columns(cols.data(), q, ny, ncol, col);
            const uint32_t j = compact_index(col[1][1], z);
            if ((int64_t)j != rank[c]) {
              printf("own number mismatch at (%d,%d,%d): %u vs %lld\n", x, y, z, j, (long long)rank[c]);
              return 1;
            }
            uint32_t J[19];
            neighbor_numbers(col, j, z, J);
            for (int k = 1; k < 19; ++k) {
              const int sx = x + ex_of(k), sy = y + ey_of(k), sz = z + ez_of(k);
              if (sx < 0 || sx >= nx || sy < 0 || sy >= ny || sz < 0 || sz >= nz) continue;
              const int64_t cn = ((int64_t)sx * ny + sy) * nz + sz;
              if (mask[cn] == BC_SOLID) continue;
              if ((int64_t)J[k] != rank[cn]) {
                printf("neighbour %d of (%d,%d,%d) in %dx%dx%d: %u vs %lld\n", k, x, y, z, nx, ny, nz, J[k],
                       (long long)rank[cn]);
                return 1;
              }
            }
          }
      ++cases;
    }
  printf("ok %d\n", cases);
  return 0;
}
