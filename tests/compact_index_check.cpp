// CPU check of the compact fluid addressing used by the HIP pass
// (xadvcool_amd/csrc/lattice.h), on random masks:
//   * own_offset() of every non-solid cell is its rank among non-solid cells
//     in (x, y, z) order times the value size;
//   * neighbour_offsets() gives, for every neighbour c + e_k inside the lattice
//     and non-solid, that neighbour's rank times the value size, plus one
//     plane (NFB bytes) for even k;
//   * every column entry read around a cell lies inside the padded table;
//   * div_nz() equals c / nz for nz = 1..32, exhaustively for c < 2^21 and at
//     every multiple of nz (and its neighbours) up to 2^26.
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

template <uint32_t RS>
int check_mask(const std::vector<int8_t>& mask, int nx, int ny, int nz) {
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<ColInfo> table((size_t)padded_columns(nx, ny));
  const int64_t nf = build_columns(mask.data(), nx, ny, nz, table.data());
  const ColInfo* cols = table.data() + ny + 1;
  std::vector<int64_t> rank(n, -1);
  int64_t r = 0;
  for (int64_t c = 0; c < n; ++c)
    if (mask[c] != BC_SOLID) rank[c] = r++;
  if (r != nf) {
    printf("count mismatch %lld vs %lld\n", (long long)r, (long long)nf);
    return 1;
  }
  const uint32_t NFB = (uint32_t)((nf + 63) / 64 * 64) * RS;
  const uint32_t magic = nz_magic((uint32_t)nz);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t c = ((int64_t)x * ny + y) * nz + z;
        if (mask[c] == BC_SOLID) continue;
        const uint32_t q = div_nz((uint32_t)c, (uint32_t)nz, magic);
        if (q != (uint32_t)(x * ny + y) || (int)((uint32_t)c - q * nz) != z) {
          printf("div_nz wrong for c=%lld nz=%d\n", (long long)c, nz);
          return 1;
        }
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy) {
            const int64_t e = ny + 1 + (int64_t)q + dx * ny + dy;
            if (e < 0 || e >= padded_columns(nx, ny)) {
              printf("column entry %lld outside the padded table\n", (long long)e);
              return 1;
            }
          }
        const uint32_t own = own_offset<RS>(cols, q, z);
        uint32_t j4 = 0, JB[19];
        neighbour_offsets<RS>(cols, q, z, ny, NFB, j4, JB);
        if ((int64_t)own != rank[c] * RS || j4 != own) {
          printf("own offset mismatch at (%d,%d,%d): %u / %u vs %lld\n", x, y, z, own, j4, (long long)(rank[c] * RS));
          return 1;
        }
        for (int k = 1; k < 19; ++k) {
          const int sx = x + ex_of(k), sy = y + ey_of(k), sz = z + ez_of(k);
          if (sx < 0 || sx >= nx || sy < 0 || sy >= ny || sz < 0 || sz >= nz) continue;
          const int64_t cn = ((int64_t)sx * ny + sy) * nz + sz;
          if (mask[cn] == BC_SOLID) continue;
          const int64_t want = rank[cn] * RS + ((k & 1) ? 0 : NFB);
          if ((int64_t)JB[k] != want) {
            printf("neighbour %d of (%d,%d,%d) in %dx%dx%d: %u vs %lld\n", k, x, y, z, nx, ny, nz, JB[k],
                   (long long)want);
            return 1;
          }
        }
      }
  return 0;
}

int main() {
  std::mt19937 rng(12345);
  const int shapes[][3] = {{3, 3, 1}, {4, 5, 2}, {7, 3, 32}, {9, 8, 31}, {12, 11, 16}, {5, 17, 32}, {3, 1, 5}, {6, 2, 17}};
  int cases = 0;
  for (const auto& sh : shapes)
    for (double solid : {0.0, 0.3, 0.7, 0.95, 1.0}) {
      const int nx = sh[0], ny = sh[1], nz = sh[2];
      std::vector<int8_t> mask((size_t)nx * ny * nz);
      std::uniform_real_distribution<double> U(0.0, 1.0);
      for (auto& m : mask) m = U(rng) < solid ? BC_SOLID : (int8_t)(U(rng) < 0.5 ? BC_FLUID : BC_OUTLET);
      if (check_mask<4>(mask, nx, ny, nz) || check_mask<8>(mask, nx, ny, nz)) return 1;
      ++cases;
    }
  for (uint32_t nz = 1; nz <= 32; ++nz) {
    const uint32_t m = nz_magic(nz);
    for (uint32_t c = 0; c < (1u << 21); ++c)
      if (div_nz(c, nz, m) != c / nz) {
        printf("div_nz(%u, %u) = %u\n", c, nz, div_nz(c, nz, m));
        return 1;
      }
    for (uint32_t t = 0; t <= (1u << 26) / nz; t += 1 + t / 1024)
      for (int d = -1; d <= 1; ++d) {
        const int64_t c = (int64_t)t * nz + d;
        if (c < 0 || c >= (1 << 26)) continue;
        if (div_nz((uint32_t)c, nz, m) != (uint32_t)c / nz) {
          printf("div_nz(%lld, %u) = %u\n", (long long)c, nz, div_nz((uint32_t)c, nz, m));
          return 1;
        }
      }
    ++cases;
  }
  printf("ok %d\n", cases);
  return 0;
}
