// Index arithmetic of the tessellation tiles and work items, shared by the
// device kernels and a host-side brute-force check (compiled by hipcc for
// the device and by a plain C++ compiler for the CPU tests).
//
// A tile is a run of consecutive cells of one z-plane: `rows` whole rows of
// cells, or, when a row holds more than kTileCells cells, one of the `segs`
// segments of a row (then rows = 1).  Tiles are numbered in cell order,
// tile = k * cols + col with col = (j / rows) * segs + i / seg_w, so their
// triangle counts scan into the serial order of the reference
// (contour/main.c:49-112: k, then j, then i).  A work item is one column
// `col` over zc consecutive planes, numbered column-fastest:
// item = (k0 / zc) * cols + col.
#pragma once

#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#define C3D_TD __host__ __device__ __forceinline__
#else
#define C3D_TD inline
#endif

namespace c3d {

constexpr int kTileCells = 2048;     // cells of a tessellation tile at most
constexpr int kDefaultZPlanes = 4;   // z-planes a work item marches by default (2-6, 8, 16 measured at 512^3: 4 best)

// Division by a runtime constant with a multiply-high (Granlund-Montgomery,
// "round-up" variant): q = (umulhi(n, m) + n) >> s for every 32-bit n.
struct FastDiv {
  uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv r;
  r.d = d;
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) s++;
  r.s = s;
  r.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
  return r;
}

C3D_TD uint32_t fdiv(uint32_t n, const FastDiv& f) {
  const uint64_t hi = (uint64_t(n) * f.m) >> 32;
  return uint32_t((hi + n) >> f.s);
}

struct TileGeo {
  uint32_t rows;   // rows of cells per tile (1 when rows are split)
  uint32_t segs;   // segments per row
  uint32_t seg_w;  // cells per row segment (ncx when segs == 1)
  uint32_t cols;   // tiles per plane
  uint32_t zc;     // planes per work item
  uint32_t nzc;    // work items per column, ceil(ncz / zc)
  FastDiv div_cols, div_seg_w, div_rows;
};

// The tile layout of an ncx x ncy x ncz cell lattice: tiles of up to
// kTileCells cells, at most max_rows rows each (0: as many as fit), work
// items of zc planes (0: kDefaultZPlanes).
inline TileGeo make_tile_geo(uint32_t ncx, uint32_t ncy, uint32_t ncz, uint32_t max_rows, uint32_t zc) {
  TileGeo g{};
  if (ncx <= uint32_t(kTileCells)) {
    uint32_t r = uint32_t(kTileCells) / ncx;
    if (max_rows && r > max_rows) r = max_rows;
    if (r > ncy) r = ncy;
    g.rows = r < 1 ? 1 : r;
    g.segs = 1;
    g.seg_w = ncx;
  } else {
    g.rows = 1;
    g.segs = (ncx + kTileCells - 1) / kTileCells;
    g.seg_w = (ncx + g.segs - 1) / g.segs;  // balanced segments of at most kTileCells
  }
  g.cols = (ncy + g.rows - 1) / g.rows * g.segs;
  g.zc = zc ? zc : uint32_t(kDefaultZPlanes);
  if (g.zc > ncz) g.zc = ncz;
  g.nzc = (ncz + g.zc - 1) / g.zc;
  g.div_cols = make_fastdiv(g.cols);
  g.div_seg_w = make_fastdiv(g.seg_w);
  g.div_rows = make_fastdiv(g.rows);
  return g;
}

// The cells of tile (k, col): a contiguous run of n cells from cell index
// c0, w cells per row; cell p of the tile has its corner 0 at lattice index
// l0 + tile_corner_off(p).
struct TileSpan {
  uint32_t j0, i0, w, n;
  uint32_t rows;  // rows of cells (n = rows * w)
  uint32_t c0;
  uint64_t l0;
};

C3D_TD TileSpan tile_span(const TileGeo& g, uint32_t ncx, uint32_t ncy, uint32_t nx, uint32_t ny, uint32_t k,
                          uint32_t col) {
  TileSpan t;
  const uint32_t jb = g.segs == 1 ? col : col / g.segs;
  const uint32_t is = col - jb * g.segs;
  t.j0 = jb * g.rows;
  t.i0 = is * g.seg_w;
  const uint32_t rows = g.rows < ncy - t.j0 ? g.rows : ncy - t.j0;
  t.w = g.segs == 1 ? ncx : (g.seg_w < ncx - t.i0 ? g.seg_w : ncx - t.i0);
  t.n = rows * t.w;
  t.rows = rows;
  t.c0 = (k * ncy + t.j0) * ncx + t.i0;
  t.l0 = (uint64_t(k) * ny + t.j0) * nx + t.i0;
  return t;
}

// lattice offset of tile cell p's corner 0 from the tile's (the same at
// every plane of a column); div_ncx divides by ncx
C3D_TD uint32_t tile_corner_off(const TileGeo& g, uint32_t ncx, uint32_t nx, const FastDiv& div_ncx, uint32_t p) {
  if (g.segs != 1) return p;  // one row segment
  const uint32_t jj = fdiv(p, div_ncx);
  return jj * nx + (p - jj * ncx);
}

// a work item's column and planes [k0, k1)
C3D_TD void item_range(const TileGeo& g, uint32_t ncz, uint32_t item, uint32_t& col, uint32_t& k0, uint32_t& k1) {
  const uint32_t z = fdiv(item, g.div_cols);
  col = item - z * g.cols;
  k0 = z * g.zc;
  k1 = k0 + g.zc < ncz ? k0 + g.zc : ncz;
}

// the tile holding cell c
C3D_TD uint32_t tile_of_cell(const TileGeo& g, uint32_t ncx, uint32_t ncy, const FastDiv& div_ncx,
                             const FastDiv& div_ncy, uint32_t c) {
  const uint32_t jk = fdiv(c, div_ncx);
  const uint32_t i = c - jk * ncx;
  const uint32_t k = fdiv(jk, div_ncy);
  const uint32_t j = jk - k * ncy;
  const uint32_t jb = g.rows == 1 ? j : fdiv(j, g.div_rows);
  const uint32_t is = g.segs == 1 ? 0u : fdiv(i, g.div_seg_w);
  return k * g.cols + jb * g.segs + is;
}

}  // namespace c3d
