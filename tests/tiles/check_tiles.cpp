// Brute-force check of the tessellation tile and work-item arithmetic of
// c3d_tiles.h (the header the device kernels use).  For each lattice given
// as "ncx ncy ncz max_rows zc" on the command line, walks every work item,
// plane and tile cell as the kernels do and checks that
//   - every cell is visited exactly once, at index c0 + (k - k0) * ncx * ncy + p;
//   - tiles hold at most kTileCells cells, are contiguous runs numbered in
//     cell order (tile = k * cols + col), and tile_of_cell inverts them;
//   - the corner-0 lattice index the kernels form equals (k * ny + j) * nx + i,
//     so every corner read (+1, +nx, +nx*ny) stays inside the field;
//   - work items cover the planes of every column once.
// Prints one line per lattice; exits 1 on the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "c3d_tiles.h"

using namespace c3d;

static int fail(const char* what, unsigned long long a, unsigned long long b) {
  std::printf("FAIL %s (%llu, %llu)\n", what, a, b);
  return 1;
}

static int check(uint32_t ncx, uint32_t ncy, uint32_t ncz, uint32_t max_rows, uint32_t zc) {
  const uint32_t nx = ncx + 1, ny = ncy + 1, nz = ncz + 1;
  const uint64_t ncells = uint64_t(ncx) * ncy * ncz;
  const TileGeo g = make_tile_geo(ncx, ncy, ncz, max_rows, zc);
  const FastDiv dx = make_fastdiv(ncx), dy = make_fastdiv(ncy);
  std::vector<uint8_t> seen(ncells, 0);
  std::vector<uint32_t> tile_first(uint64_t(ncz) * g.cols, 0xffffffffu), tile_n(uint64_t(ncz) * g.cols, 0);
  const uint32_t nitems = g.cols * g.nzc;
  for (uint32_t item = 0; item < nitems; item++) {
    uint32_t col, k0, k1;
    item_range(g, ncz, item, col, k0, k1);
    if (col >= g.cols || k0 >= k1 || k1 > ncz) return fail("item range", item, k0);
    const TileSpan t = tile_span(g, ncx, ncy, nx, ny, k0, col);
    if (t.n == 0 || t.n > uint32_t(kTileCells)) return fail("tile size", item, t.n);
    if (t.rows * t.w != t.n) return fail("tile rows", t.rows, t.n);
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t tile = k * g.cols + col;
      const uint32_t ctile = t.c0 + (k - k0) * (ncx * ncy);
      tile_first[tile] = ctile;
      tile_n[tile] = t.n;
      for (uint32_t p = 0; p < t.n; p++) {
        const uint64_t c = uint64_t(ctile) + p;
        if (c >= ncells) return fail("cell out of range", c, ncells);
        if (seen[c]++) return fail("cell visited twice", c, item);
        const uint32_t i = uint32_t(c % ncx), j = uint32_t((c / ncx) % ncy), kk = uint32_t(c / (uint64_t(ncx) * ncy));
        if (kk != k) return fail("plane", c, k);
        if (tile_of_cell(g, ncx, ncy, dx, dy, uint32_t(c)) != tile) return fail("tile_of_cell", c, tile);
        const uint64_t l = t.l0 + uint64_t(k - k0) * nx * ny + tile_corner_off(g, ncx, nx, dx, p);
        if (l != (uint64_t(k) * ny + j) * nx + i) return fail("corner index", c, l);
        if (l + uint64_t(nx) * ny + nx + 1 >= uint64_t(nx) * ny * nz) return fail("corner read past the field", c, l);
      }
    }
  }
  for (uint64_t c = 0; c < ncells; c++)
    if (!seen[c]) return fail("cell never visited", c, 0);
  for (uint64_t tl = 0; tl + 1 < tile_first.size(); tl++)
    if (tile_first[tl] + tile_n[tl] != tile_first[tl + 1]) return fail("tiles not consecutive", tl, tile_first[tl]);
  if (tile_first[0] != 0) return fail("first tile", tile_first[0], 0);
  std::printf("ok %u x %u x %u rows<=%u zc=%u: rows %u segs %u seg_w %u cols %u zc %u items %u\n", ncx, ncy, ncz,
              max_rows, zc, g.rows, g.segs, g.seg_w, g.cols, g.zc, nitems);
  return 0;
}

int main(int argc, char** argv) {
  for (int a = 1; a + 4 < argc; a += 5) {
    const uint32_t v[5] = {uint32_t(std::atoi(argv[a])), uint32_t(std::atoi(argv[a + 1])),
                           uint32_t(std::atoi(argv[a + 2])), uint32_t(std::atoi(argv[a + 3])),
                           uint32_t(std::atoi(argv[a + 4]))};
    if (check(v[0], v[1], v[2], v[3], v[4])) return 1;
  }
  // the multiply-high division against plain division
  const uint32_t ds[] = {1, 2, 3, 7, 63, 511, 1023, 2047, 2048, 2049, 65535, 65536, 1000003, 0x7fffffffu, 0xffffffffu};
  uint64_t x = 88172645463325252ull;
  for (uint32_t d : ds) {
    const FastDiv f = make_fastdiv(d);
    for (int r = 0; r < 200000; r++) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      const uint32_t n = r < 4 ? (r == 0 ? 0u : r == 1 ? 0xffffffffu : r == 2 ? d : d - 1) : uint32_t(x);
      if (fdiv(n, f) != n / d) return fail("fdiv", n, d);
    }
  }
  std::printf("ok fdiv\n");
  return 0;
}
