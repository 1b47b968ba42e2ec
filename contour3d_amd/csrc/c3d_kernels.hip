// Device kernels of the tessellate -> smooth -> render path for gfx950.
//
// FP64 form: every floating-point step is the reference's, in the same
// order, built with -ffp-contract=off so no multiply-add is fused; the
// results are bit-identical to the reference's FP64 arithmetic.
// FP32 form (R = float): the corner classification and the 1e-8 edge test
// stay in FP64 (topology is exact); interpolation, normals, projection,
// depth and shading run in FP32.
#include "c3d_internal.h"

#include <cfloat>
#include <type_traits>

namespace c3d {

// ----------------------------------------------------------------------
// small helpers

// ----------------------------------------------------------------------
// block-wide exclusive scan of one value per thread (256 threads)

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint32_t off = 0;
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) {
    const uint32_t s = s_wave[w];
    off += w < wave ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// ----------------------------------------------------------------------
// field access

// Field and grid pointers reach the kernels through device-side job tables,
// where the compiler cannot tell that they address global memory; read
// through gptr they become global loads (scalar base + vector offset)
// instead of flat ones, which also wait on the LDS counter.
#define C3D_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const C3D_GLOBAL T* gptr(const T* p) {
  return (const C3D_GLOBAL T*)p;
}

struct CellLoc {
  uint32_t i, j, k;
  uint64_t p0;  // lattice index of corner 0
};

__device__ __forceinline__ CellLoc locate(const FieldDev& f, uint32_t c) {
  CellLoc L;
  const uint32_t jk = fdiv(c, f.div_ncx);
  L.i = c - jk * f.ncx;
  L.k = fdiv(jk, f.div_ncy);
  L.j = jk - L.k * f.ncy;
  L.p0 = (uint64_t(L.k) * f.ny + L.j) * f.nx + L.i;
  return L;
}

__device__ __forceinline__ uint64_t corner_offset(const FieldDev& f, uint32_t c) {
  // selects rather than 64-bit products: c is a corner number (0-7)
  const uint64_t nxy = uint64_t(f.nx) * f.ny;
  return uint64_t(c & 1u) + uint64_t(c & 2u ? f.nx : 0u) + (c & 4u ? nxy : 0ull);
}

template <typename R>
__device__ __forceinline__ V3<R> corner_pos(const FieldDev& f, const CellLoc& L, uint32_t c) {
  if (f.cx) {
    const uint64_t p = L.p0 + corner_offset(f, c);
    return {R(gptr(f.cx)[p]), R(gptr(f.cy)[p]), R(gptr(f.cz)[p])};
  }
  return {R(gptr(f.gx)[L.i + (c & 1u)]), R(gptr(f.gy)[L.j + ((c >> 1) & 1u)]), R(gptr(f.gz)[L.k + (c >> 2)])};
}

// ----------------------------------------------------------------------
// block-wide exclusive scan of a 64-bit value per thread (256 threads); used
// on four 16-bit counters packed in one word, one per threshold

__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* s_wave, uint64_t& total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) {
    const uint64_t s = s_wave[w];
    off += w < wave ? s : 0ull;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// the eight corner values of cell c (corner q = i + 2j + 4k offsets)
__device__ __forceinline__ void load_corners(const FieldDev& f, const CellLoc& L, double v[8]) {
  const C3D_GLOBAL double* p = gptr(f.f) + L.p0;
  const uint64_t sy = f.nx, sz = uint64_t(f.nx) * f.ny;
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[sy];
  v[3] = p[sy + 1];
  v[4] = p[sz];
  v[5] = p[sz + 1];
  v[6] = p[sz + sy];
  v[7] = p[sz + sy + 1];
}

// Topology tables, built once by upload_tables from the tetra and case
// tables of triangulate.c:145-152, 204-270:
//   c_ntri[code]            triangles of a cell with that 8-bit corner code
//   c_cslot[code * 12 + s]  triangle slot s of such a cell: bits 0-17 hold the
//                           (tail, head) cube corners of its vertices 0, 1, 2
//                           (3 + 3 bits each, reversal of triangulate.c:108-112
//                           already applied), bits 18-32 the three 5-bit
//                           cell-local edge ids (triangulate.c:23-32)
__constant__ uint8_t c_ntri[256];
__constant__ uint64_t c_cslot[256 * 12];
//   c_emask[code * 19 + e]  the slots (bit s) of such a cell whose triangle
//                           has a vertex on cell-local edge e
__constant__ uint16_t c_emask[256 * 19];

__device__ __forceinline__ uint32_t corner_code(double thr, const double v[8]) {
  uint32_t code = 0;
#pragma unroll
  for (int n = 0; n < 8; n++) code |= uint32_t(thr < v[n]) << n;
  return code;
}

// triangle counts of up to four thresholds packed 16 bits apart, from an
// LDS copy of c_ntri
__device__ __forceinline__ uint64_t packed_counts(const ThrSet& ts, const double v[8], uint32_t code[kMaxThr],
                                                  const uint8_t* s_ntri) {
  uint64_t packed = 0;
#pragma unroll
  for (int q = 0; q < kMaxThr; q++) {
    uint32_t cd = 0;
    if (q < ts.n) cd = corner_code(ts.thr[q], v);
    code[q] = cd;
    packed |= uint64_t(s_ntri[cd]) << (16 * q);
  }
  return packed;
}

// ----------------------------------------------------------------------
// tessellation tiles and work items

// the tile arithmetic of c3d_tiles.h on a field
__device__ __forceinline__ TileSpan tile_span(const FieldDev& f, uint32_t k, uint32_t col) {
  return tile_span(f.tg, f.ncx, f.ncy, f.nx, f.ny, k, col);
}

__device__ __forceinline__ uint32_t tile_corner_off(const FieldDev& f, const TileSpan&, uint32_t p) {
  return tile_corner_off(f.tg, f.ncx, f.nx, f.div_ncx, p);
}

__device__ __forceinline__ void item_range(const FieldDev& f, uint32_t item, uint32_t& col, uint32_t& k0,
                                           uint32_t& k1) {
  item_range(f.tg, f.ncz, item, col, k0, k1);
}

// the tile holding cell c of an object's field (the stored mesh's lookup)
__device__ __forceinline__ uint32_t tile_of_cell(const ObjMeta& om, uint32_t c) {
  return tile_of_cell(om.tg, om.ncx, om.ncy, om.div_ncx, om.div_ncy, c);
}

// The 4-bit code of four corners against threshold thr: bit n = thr < v[n]
// (triangulate.c:176-179).  A cell's 8-bit corner code is the code of its
// lower plane (corners 0-3) | the code of its upper plane (corners 4-7) << 4,
// and a cell's upper code is the lower code of the cell above it.
__device__ __forceinline__ uint32_t plane_code(double thr, const double v[4]) {
  uint32_t c = 0;
#pragma unroll
  for (int n = 0; n < 4; n++) c |= uint32_t(thr < v[n]) << n;
  return c;
}

// the four corners of one plane of a cell, from its corner 0 at p
__device__ __forceinline__ void load_plane4(const C3D_GLOBAL double* p, uint32_t nx, double v[4]) {
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[nx];
  v[3] = p[nx + 1];
}

// plane codes of up to four thresholds, 4 bits apart
__device__ __forceinline__ uint32_t plane_codes(const ThrSet& ts, const double v[4]) {
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < kMaxThr; q++)
    if (q < ts.n) c |= plane_code(ts.thr[q], v) << (4 * q);
  return c;
}

// ----------------------------------------------------------------------
// the walk over work items
//
// Workgroups are dispatched round-robin over the 8 XCDs.  The work items of
// each field pass are split into 8 contiguous segments, one per XCD, so that
// every XCD gets an equal share of every pass, however unequal the passes of
// a batch are.  Work items are numbered column-fastest within a z-range, so
// the workgroups an XCD runs at once march neighbouring columns over the
// same planes (the rows they share meet in that XCD's L2).  XCD x walks
// positions 0 .. end-1: position p lies in the pass jb with
// jobs[jb].xoff <= p < jobs[jb + 1].xoff, at item x * seg + p - xoff of it
// (seg = ceil(items of the pass / 8)); the XCD's workgroups stride through
// the positions.  The grid must be a multiple of 8.
struct XcdWalk {
  uint32_t x, begin, end, step;
};

__device__ __forceinline__ XcdWalk xcd_walk(const TessJob* __restrict__ jobs, int njobs, uint32_t nitems) {
  const TessJob& last = jobs[njobs - 1];
  XcdWalk w;
  w.x = blockIdx.x % 8u;
  w.begin = blockIdx.x / 8u;
  w.end = last.xoff + (nitems - last.item0 + 7u) / 8u;
  w.step = gridDim.x / 8u;
  return w;
}

// The pass and its work item at position p of XCD x's walk; false where the
// XCD's segment of that pass is shorter than seg (the last ones).
__device__ __forceinline__ bool xcd_item(const TessJob* __restrict__ jobs, int njobs, uint32_t nitems, uint32_t x,
                                         uint32_t p, uint32_t& jb, uint32_t& item) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].xoff <= p) lo = mid;
    else hi = mid - 1;
  }
  jb = uint32_t(lo);
  const uint32_t t0 = jobs[lo].item0;
  const uint32_t nt = (lo + 1 < njobs ? jobs[lo + 1].item0 : nitems) - t0;
  item = x * ((nt + 7u) / 8u) + (p - jobs[lo].xoff);
  return item < nt;
}

// ----------------------------------------------------------------------
// tessellation: count pass
//
// Per work item, each thread takes column segments of R rows of the tile
// (R = 4, 2 or 1, the largest that divides the tile's rows): a segment's
// R cells share their rows of lattice points, so a plane costs 2 (R + 1)
// point loads and threshold compares for R cells instead of 4 R.  The
// cells' lower-plane codes (4 bits per threshold) stay in registers while
// the item marches up its column, so the field is read about once.  Only
// the tile totals are written, so the cells may be taken in any order.

// the bits of one lattice point: thr_q < v at bit 4q
__device__ __forceinline__ uint32_t point_bits(const ThrSet& ts, double v) {
  uint32_t b = 0;
#pragma unroll
  for (int q = 0; q < kMaxThr; q++)
    if (q < ts.n) b |= uint32_t(ts.thr[q] < v) << (4 * q);
  return b;
}

// the plane codes (plane_codes order) of the R cells of a segment whose
// top-left point is at p: points of rows 0..R, columns 0 and 1
template <int R>
__device__ __forceinline__ void segment_codes(const ThrSet& ts, const C3D_GLOBAL double* p, uint32_t nx,
                                              uint32_t out[R]) {
  double v[R + 1][2];
#pragma unroll
  for (int jj = 0; jj <= R; jj++) {
    v[jj][0] = p[uint64_t(jj) * nx];
    v[jj][1] = p[uint64_t(jj) * nx + 1];
  }
  uint32_t b0 = point_bits(ts, v[0][0]) | point_bits(ts, v[0][1]) << 1;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const uint32_t b1 = point_bits(ts, v[r + 1][0]) | point_bits(ts, v[r + 1][1]) << 1;
    out[r] = b0 | b1 << 2;
    b0 = b1;
  }
}

template <int R>
__device__ __forceinline__ void count_item(const FieldDev& f, const ThrSet& ts, const TileSpan& t, uint32_t col,
                                           uint32_t k0, uint32_t k1, const uint8_t* s_ntri, uint64_t* s_wave,
                                           uint32_t* __restrict__ cnt) {
  constexpr int M = kTileChunks / R;  // segments per thread at most (R * M = kTileCells / kThreads cells)
  const uint32_t nseg = (t.rows / R) * t.w;
  const uint64_t nxy = uint64_t(f.nx) * f.ny;
  uint32_t off[M], low[M][R];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const uint32_t sg = m * kThreads + threadIdx.x;
    off[m] = 0;
#pragma unroll
    for (int r = 0; r < R; r++) low[m][r] = 0;
    if (sg < nseg) {
      // segment sg: rows R * (sg / w) .. + R - 1, column sg % w of the tile
      const uint32_t rs = R == 1 && f.tg.segs != 1 ? 0u : fdiv(sg, f.div_ncx);
      off[m] = rs * R * f.nx + (sg - rs * t.w);
      segment_codes<R>(ts, gptr(f.f) + t.l0 + off[m], f.nx, low[m]);
    }
  }
  for (uint32_t k = k0; k < k1; k++) {
    const C3D_GLOBAL double* up = gptr(f.f) + t.l0 + uint64_t(k - k0 + 1) * nxy;
    uint64_t acc = 0;
#pragma unroll
    for (int m = 0; m < M; m++) {
      const uint32_t sg = m * kThreads + threadIdx.x;
      if (sg < nseg) {
        uint32_t hi[R];
        segment_codes<R>(ts, up + off[m], f.nx, hi);
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
          for (int q = 0; q < kMaxThr; q++) {
            const uint32_t cd = ((low[m][r] >> (4 * q)) & 15u) | ((hi[r] >> (4 * q)) & 15u) << 4;
            acc += uint64_t(q < ts.n ? s_ntri[cd] : 0u) << (16 * q);
          }
          low[m][r] = hi[r];
        }
      }
    }
    uint64_t tot;
    block_excl_scan64(acc, s_wave, tot);
    const uint32_t tile = k * f.tg.cols + col;
    if (threadIdx.x < uint32_t(ts.n)) cnt[ts.boff[threadIdx.x] + tile] = uint32_t(tot >> (16 * threadIdx.x)) & 0xffffu;
  }
}

// The plane codes of every cell of tile t on the plane whose tile corner is
// at p, by column segments of R rows as in count_item, into out[cell]
// (cell = row * w + column within the tile).
template <int R>
__device__ __forceinline__ void classify_plane_r(const FieldDev& f, const ThrSet& ts, const TileSpan& t,
                                                 const C3D_GLOBAL double* p, uint16_t* out) {
  constexpr int M = kTileChunks / R;
  const uint32_t nseg = (t.rows / R) * t.w;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const uint32_t sg = m * kThreads + threadIdx.x;
    if (sg < nseg) {
      const uint32_t rs = R == 1 && f.tg.segs != 1 ? 0u : fdiv(sg, f.div_ncx);
      const uint32_t c = sg - rs * t.w;
      uint32_t codes[R];
      segment_codes<R>(ts, p + rs * R * f.nx + c, f.nx, codes);
#pragma unroll
      for (int r = 0; r < R; r++) out[(rs * R + r) * t.w + c] = uint16_t(codes[r]);
    }
  }
}

__device__ __forceinline__ void classify_plane(const FieldDev& f, const ThrSet& ts, const TileSpan& t,
                                               const C3D_GLOBAL double* p, uint16_t* out) {
  if (t.rows % 4u == 0) classify_plane_r<4>(f, ts, t, p, out);
  else if (t.rows % 2u == 0) classify_plane_r<2>(f, ts, t, p, out);
  else classify_plane_r<1>(f, ts, t, p, out);
}

// seven waves per SIMD (72 VGPRs, no spill): 0.265 ms against 0.282 at six at
// 512^3 x 3 thresholds; eight spill and take 0.41 ms
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7))) void k_count(
    const TessJob* __restrict__ jobs, int njobs, uint32_t nitems, uint32_t* __restrict__ cnt) {
  __shared__ uint64_t s_wave[kThreads / 64];
  __shared__ uint8_t s_ntri[256];
  s_ntri[threadIdx.x] = c_ntri[threadIdx.x];
  __syncthreads();
  const XcdWalk W = xcd_walk(jobs, njobs, nitems);
  for (uint32_t pp = W.begin; pp < W.end; pp += W.step) {
    uint32_t jb, item;
    if (!xcd_item(jobs, njobs, nitems, W.x, pp, jb, item)) continue;
    const FieldDev f = jobs[jb].f;
    const ThrSet ts = jobs[jb].ts;
    uint32_t col, k0, k1;
    item_range(f, item, col, k0, k1);
    const TileSpan t = tile_span(f, k0, col);
    if (t.rows % 4u == 0) count_item<4>(f, ts, t, col, k0, k1, s_ntri, s_wave, cnt);
    else if (t.rows % 2u == 0) count_item<2>(f, ts, t, col, k0, k1, s_ntri, s_wave, cnt);
    else count_item<1>(f, ts, t, col, k0, k1, s_ntri, s_wave, cnt);
  }
}

// ----------------------------------------------------------------------
// scan of tile counts into triangle starts (u64, exclusive, start[n] = total)


__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ cnt, uint64_t n,
                                                          uint64_t* __restrict__ tile_sum) {
  __shared__ uint64_t s[kThreads];
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
  uint64_t acc = 0;
  for (int q = 0; q < kScanTile / kThreads; q++) {
    const uint64_t i = base + q * kThreads + threadIdx.x;
    if (i < n) acc += cnt[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = s[0];
}

// single workgroup: exclusive scan of the tile sums in place
__global__ __launch_bounds__(kThreads) void k_scan_tiles(uint64_t* __restrict__ tile_sum, uint64_t ntile) {
  __shared__ uint64_t s[kThreads];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < ntile; b += kThreads) {
    const uint64_t i = b + threadIdx.x;
    const uint64_t v = i < ntile ? tile_sum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t d = 1; d < kThreads; d <<= 1) {
      const uint64_t y = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < ntile) tile_sum[i] = carry + s[threadIdx.x] - v;
    const uint64_t tot = s[kThreads - 1];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sum[ntile] = carry;
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t* __restrict__ cnt, uint64_t n,
                                                         const uint64_t* __restrict__ tile_off,
                                                         uint64_t* __restrict__ start) {
  __shared__ uint32_t s_wave[kThreads / 64];
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
  uint64_t carry = tile_off[blockIdx.x];
  for (int q = 0; q < kScanTile / kThreads; q++) {
    const uint64_t i = base + q * kThreads + threadIdx.x;
    const uint32_t v = i < n ? cnt[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, s_wave, tot);
    if (i < n) start[i] = carry + ex;
    carry += tot;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[n] = tile_off[gridDim.x];
}

// Object bases and counts, the frame total and each image's triangle range,
// read off the tile starts on the device (so that no host round trip sits
// between counting and emitting when a frame is replayed from a graph).
template <typename R>
__global__ void k_plan(const uint64_t* __restrict__ start, uint64_t ntiles, ObjMeta* __restrict__ meta, int nobj,
                       RasterJob<R>* __restrict__ rjobs, int njobs, FrameHdr* __restrict__ hdr, uint64_t cap) {
  // objects are never empty (at least one cell), so nobj > 0 implies ntiles > 0
  for (int o = threadIdx.x; o < nobj; o += blockDim.x) {
    const uint64_t b = start[meta[o].tile_off];
    const uint64_t e = o + 1 < nobj ? start[meta[o + 1].tile_off] : start[ntiles];
    meta[o].base = b;
    meta[o].count = e - b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t ntri = ntiles ? start[ntiles] : 0;
    uint64_t wk = 0;
    for (int j = 0; j < njobs; j++) {
      const uint64_t t0 = rjobs[j].obj0 < uint32_t(nobj) ? meta[rjobs[j].obj0].base : ntri;
      const uint64_t t1 = rjobs[j].obj1 < uint32_t(nobj) ? meta[rjobs[j].obj1].base : ntri;
      // each image's work starts on a workgroup boundary, so that every
      // workgroup round of the raster kernels belongs to one image
      wk = (wk + kThreads - 1) / kThreads * kThreads;
      rjobs[j].tri0 = t0;
      rjobs[j].tri1 = t1;
      rjobs[j].wk0 = wk;
      wk += t1 - t0;
    }
    hdr->ntri = ntri;
    hdr->nwork = wk;
    hdr->nfrag = 0;
    // triangles take ids 1..ntri and line fragments the ids after them, all
    // in 32 bits; a frame beyond that is flagged like one beyond `cap` (the
    // host then rejects it after its read back)
    hdr->overflow = (ntri > cap || !frame_ids_fit(ntri, hdr->nline)) ? 1u : 0u;
  }
}

// linear interpolation along a tetra edge (triangulate.c:40-73); the 1e-8
// test is on the FP64 corner values in both forms
template <typename R>
__device__ __forceinline__ V3<R> interp(double thr, V3<R> p0, double v0, V3<R> p1, double v1) {
  if (fabs(v1 - v0) < 1.e-8) return p0;
  R fct;
  if constexpr (sizeof(R) == 8) {
    fct = (thr - v0) / (v1 - v0);
  } else {
    fct = R(thr - v0) / R(v1 - v0);
  }
  return vadd(p0, vscale(fct, vsub(p1, p0)));
}

// positions of the triangle in slot entry `ent` of cell L from its corner
// values v (the arithmetic of the emit pass, from registers)
template <typename R>
__device__ __forceinline__ void slot_positions(const FieldDev& f, const CellLoc& L, double thr, uint64_t ent,
                                               const double v[8], V3<R> P[3]) {
#pragma unroll
  for (int vv = 0; vv < 3; vv++) {
    const uint32_t tail = uint32_t(ent >> (6 * vv)) & 7u;
    const uint32_t head = uint32_t(ent >> (6 * vv + 3)) & 7u;
    double vt = v[0], vh = v[0];
#pragma unroll
    for (int q = 1; q < 8; q++) {
      vt = tail == uint32_t(q) ? v[q] : vt;
      vh = head == uint32_t(q) ? v[q] : vh;
    }
    P[vv] = interp<R>(thr, corner_pos<R>(f, L, tail), vt, corner_pos<R>(f, L, head), vh);
  }
}

// The six grid coordinates of cell L (x0 x1 y0 y1 z0 z1) on a grid lattice.
__device__ __forceinline__ void cell_grid(const FieldDev& f, const CellLoc& L, double g6[6]) {
  g6[0] = gptr(f.gx)[L.i];
  g6[1] = gptr(f.gx)[L.i + 1];
  g6[2] = gptr(f.gy)[L.j];
  g6[3] = gptr(f.gy)[L.j + 1];
  g6[4] = gptr(f.gz)[L.k];
  g6[5] = gptr(f.gz)[L.k + 1];
}

// slot_positions with the cell's corner values v and, on a grid lattice,
// its six grid coordinates g6 (cell_grid) already in registers
template <typename R>
__device__ __forceinline__ void slot_positions_reg(const FieldDev& f, const CellLoc& L, double thr, uint64_t ent,
                                                   const double v[8], const double g6[6], V3<R> P[3]) {
  if (f.cx) {
    slot_positions<R>(f, L, thr, ent, v, P);
    return;
  }
#pragma unroll
  for (int vv = 0; vv < 3; vv++) {
    const uint32_t tail = uint32_t(ent >> (6 * vv)) & 7u;
    const uint32_t head = uint32_t(ent >> (6 * vv + 3)) & 7u;
    double vt = v[0], vh = v[0];
#pragma unroll
    for (int q = 1; q < 8; q++) {
      vt = tail == uint32_t(q) ? v[q] : vt;
      vh = head == uint32_t(q) ? v[q] : vh;
    }
    const V3<R> pt = {R(tail & 1u ? g6[1] : g6[0]), R(tail & 2u ? g6[3] : g6[2]), R(tail & 4u ? g6[5] : g6[4])};
    const V3<R> ph = {R(head & 1u ? g6[1] : g6[0]), R(head & 2u ? g6[3] : g6[2]), R(head & 4u ? g6[5] : g6[4])};
    P[vv] = interp<R>(thr, pt, vt, ph, vh);
  }
}

// The smoothing weight of a triangle: area * unit face normal
// (triangulate.c:75-93), formed as normal.c:39-40 forms it.  The FP32 form
// takes the cross product itself (the same vector): a triangle whose
// vertices rounded onto one point then weighs 0 instead of NaN
// (0 * (n / 0)), which would poison its neighbours' normals.
template <typename R>
__device__ __forceinline__ V3<R> smoothing_weight(const V3<R> P[3]) {
  const V3<R> n = vcross(vsub(P[1], P[0]), vsub(P[2], P[0]));
  if constexpr (sizeof(R) == 8) {
    const R area = sqrt(vdot(n, n));
    const V3<R> fn = vscale(R(1) / area, n);
    return vscale(area, fn);
  } else {
    return n;
  }
}

__device__ __forceinline__ bool interior_cell(const FieldDev& f, const CellLoc& L) {
  return L.i >= 1 && L.i + 2 <= f.ncx && L.j >= 1 && L.j + 2 <= f.ncy && L.k >= 1 && L.k + 2 <= f.ncz;
}

// ----------------------------------------------------------------------
// vertex-normal smoothing

__constant__ NbrTable c_nbr;



// the edge word of stored triangle u (in the gather record in FP64, in the
// position record in FP32)
template <typename R>
__device__ __forceinline__ uint32_t edges_of(const MeshView<R>& m, uint64_t u) {
  if constexpr (sizeof(R) == 8) return m.aux[u].edges;
  else return m.pos[u].edges;
}

template <typename R>
__device__ __forceinline__ uint64_t lower_bound_cell(const MeshAux<R>* __restrict__ aux, uint64_t lo, uint64_t hi,
                                                     uint32_t key) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (aux[mid].cell < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Unit vertex normal of a vertex on cell-local lattice edge e of cell c
// (normal.c:46-165): the sum of area * unit face normal over every triangle
// of the cells sharing the edge that has a vertex on it -- cells in
// increasing order, triangles in slot order, starting from +0 -- then
// normalised.  A cell's triangles are found from its tile's triangle range
// (triangles are stored in cell order) by binary search on the cell index.
template <typename R>
__device__ V3<R> vertex_normal(const MeshView<R>& m, const ObjMeta& om, uint32_t c, uint32_t e) {
  const int64_t sy = om.ncx, sz = int64_t(om.ncx) * om.ncy;
  V3<R> sum = {R(0), R(0), R(0)};
  const int nn = c_nbr.n[e];
  for (int q = 0; q < nn; q++) {
    const int dk = c_nbr.v[e][q][0], dj = c_nbr.v[e][q][1], di = c_nbr.v[e][q][2];
    const uint32_t want = uint32_t(c_nbr.v[e][q][3]);
    const uint32_t c2 = uint32_t(int64_t(c) + dk * sz + dj * sy + di);
    const uint32_t tl = tile_of_cell(om, c2);
    const uint64_t hi = om.bstart[tl + 1];
    for (uint64_t u = lower_bound_cell(m.aux, om.bstart[tl], hi, c2); u < hi && m.aux[u].cell == c2; u++) {
      const uint32_t eu = edges_of(m, u);
      if ((eu & 31u) == want || ((eu >> 5) & 31u) == want || ((eu >> 10) & 31u) == want)
        sum = vadd(sum, V3<R>{m.aux[u].w[0], m.aux[u].w[1], m.aux[u].w[2]});
    }
  }
  return vunit(sum);
}

// Cell starts of one object of a stored mesh, a workgroup per tile: for
// every cell c of the tile, the number of the object's triangles in the
// tile's cells before c (a tile holds at most 12 * kTileCells triangles, so
// 16 bits), from the corner codes the emit pass stored.  The triangles of
// cell c are then bstart[tile] + cstart[c] + slot, slot < c_ntri[code[c]],
// in slot order.
__global__ __launch_bounds__(kThreads) void k_cell_starts(const uint8_t* __restrict__ code,
                                                          const ObjMeta* __restrict__ meta,
                                                          uint16_t* __restrict__ cstart) {
  __shared__ uint64_t s_wave[kThreads / 64];
  __shared__ uint8_t s_ntri[256];
  s_ntri[threadIdx.x] = c_ntri[threadIdx.x];
  __syncthreads();
  const ObjMeta om = *meta;
  const uint32_t tile = blockIdx.x;
  const uint32_t k = fdiv(tile, om.tg.div_cols);
  const TileSpan t = tile_span(om.tg, om.ncx, om.ncy, om.ncx + 1u, om.ncy + 1u, k, tile - k * om.tg.cols);
  const uint8_t* cd = code + om.code_off;
  uint16_t* cs = cstart + om.code_off;
  uint32_t run = 0;
  for (uint32_t ch = 0; ch < t.n; ch += kThreads) {
    const uint32_t p = ch + threadIdx.x;
    const uint32_t n = p < t.n ? s_ntri[cd[t.c0 + p]] : 0u;
    uint64_t tot;
    const uint32_t ex = uint32_t(block_excl_scan64(n, s_wave, tot));
    if (p < t.n) cs[t.c0 + p] = uint16_t(run + ex);
    run += uint32_t(tot);
  }
}

// Vertex normals of every triangle of one object, a thread per (triangle,
// vertex): the sum of vertex_normal -- cells sharing the vertex's lattice
// edge in the neighbour table's increasing order, in each the triangles
// with a vertex on the edge in slot order (c_emask of the cell's stored
// corner code), from +0 -- gathered directly: a neighbour cell's triangles
// start at bstart[its tile] + cstart[cell].  A thread's (at most four)
// neighbour cells are fetched together.  Rows of triangles outside the
// rendered interior are NaN.
//
// (Measured against this: one 32-bit cell word of code and start, the masks
// in LDS and a persistent grid were 5% slower; a workgroup per tile, tiles
// z-fastest along each column and one contiguous run per XCD (to keep the
// planes above and below in L2), 1.55-1.66x slower: the serial order streams
// the records and the normals in address order; a workgroup per tile
// gathering a window of cells into LDS 3-4x slower.)
template <typename R>
__global__ __launch_bounds__(kThreads) void k_smooth(MeshView<R> m, const ObjMeta* __restrict__ meta,
                                                     const uint16_t* __restrict__ cstart, R* __restrict__ out) {
  const ObjMeta om = *meta;
  const uint64_t n3 = 3 * om.count;
  const uint8_t* cd = m.code + om.code_off;
  const uint16_t* cs = cstart + om.code_off;
  const int64_t sy = om.ncx, sz = int64_t(om.ncx) * om.ncy;
  const R nan = R(__longlong_as_double(0x7ff8000000000000ll));
  for (uint64_t it = uint64_t(blockIdx.x) * kThreads + threadIdx.x; it < n3; it += uint64_t(gridDim.x) * kThreads) {
    const uint64_t t = it / 3u;
    const uint32_t vv = uint32_t(it - 3u * t);
    const uint64_t u = om.base + t;
    const uint32_t c = m.aux[u].cell, ed = edges_of(m, u);
    R* o = out + 9u * t + 3u * vv;
    if (!((ed >> 15) & 1u)) {
      o[0] = nan;
      o[1] = nan;
      o[2] = nan;
      continue;
    }
    const uint32_t e = (ed >> (5u * vv)) & 31u;
    const int nn = c_nbr.n[e];
    uint32_t mask[4];
    uint64_t s2[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      mask[q] = 0;
      s2[q] = 0;
      if (q < nn) {
        // an interior cell's neighbours lie inside the lattice
        const uint32_t c2 = uint32_t(int64_t(c) + c_nbr.v[e][q][0] * sz + c_nbr.v[e][q][1] * sy + c_nbr.v[e][q][2]);
        mask[q] = c_emask[uint32_t(cd[c2]) * 19u + uint32_t(c_nbr.v[e][q][3])];
        s2[q] = om.bstart[tile_of_cell(om, c2)] + cs[c2];
      }
    }
    // each neighbour's first four hits are loaded together, then added in
    // slot order (more hits, rare, follow one at a time)
    V3<R> sum = {R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint32_t h = mask[q];
      V3<R> wv[4];
      bool hv[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint64_t u2 = s2[q] + uint32_t(__builtin_ctz(h | 0x80000000u));
        hv[r] = h != 0 && u2 < m.cap;
        wv[r] = hv[r] ? V3<R>{m.aux[u2].w[0], m.aux[u2].w[1], m.aux[u2].w[2]} : V3<R>{R(0), R(0), R(0)};
        h &= h - 1u;
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (hv[r]) sum = vadd(sum, wv[r]);
      for (; h; h &= h - 1u) {
        const uint64_t u2 = s2[q] + uint32_t(__builtin_ctz(h));
        if (u2 < m.cap) sum = vadd(sum, V3<R>{m.aux[u2].w[0], m.aux[u2].w[1], m.aux[u2].w[2]});
      }
    }
    const V3<R> nv = vunit(sum);
    o[0] = nv.x;
    o[1] = nv.y;
    o[2] = nv.z;
  }
}

// Compares object meshes built in two arithmetic forms, triangle by
// triangle (the topology is the same by construction, which this checks):
// out[0] topology mismatches (cell or edge ids), out[1] max |vertex
// difference| (as the bits of a non-negative double), out[2] max vertex
// normal angle in degrees (bits), out[3] sum of angles in 1e-9 degree
// units, out[4] interior vertices compared, out[5] / out[6] vertices whose
// angle exceeds 1 / 10 degrees.  Non-finite normals (zero-area
// neighbourhoods) are skipped.
template <typename RA, typename RB>
__global__ __launch_bounds__(kThreads) void k_mesh_compare(MeshView<RA> a, const ObjMeta* __restrict__ ma,
                                                           MeshView<RB> b, const ObjMeta* __restrict__ mb,
                                                           uint64_t count, unsigned long long* __restrict__ out) {
  const ObjMeta oa = *ma, ob = *mb;
  double verr = 0.0, amax = 0.0;
  unsigned long long asum = 0, nv = 0, bad = 0, over1 = 0, over10 = 0;
  for (uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x; t < count; t += uint64_t(gridDim.x) * kThreads) {
    const uint64_t ga = oa.base + t, gb = ob.base + t;
    const uint32_t ea = edges_of(a, ga), eb = edges_of(b, gb);
    if (ea != eb || a.aux[ga].cell != b.aux[gb].cell) {
      bad++;
      continue;
    }
#pragma unroll
    for (int c = 0; c < 9; c++) verr = fmax(verr, fabs(double(a.pos[ga].p[c]) - double(b.pos[gb].p[c])));
    if (!((ea >> 15) & 1u)) continue;
    for (int v = 0; v < 3; v++) {
      const uint32_t e = (ea >> (5 * v)) & 31u;
      const V3<RA> na = vertex_normal<RA>(a, oa, a.aux[ga].cell, e);
      const V3<RB> nb = vertex_normal<RB>(b, ob, b.aux[gb].cell, e);
      const double d = double(na.x) * double(nb.x) + double(na.y) * double(nb.y) + double(na.z) * double(nb.z);
      if (!(d == d) || !isfinite(d)) continue;
      const double ang = acos(fmin(1.0, fmax(-1.0, d))) * (180.0 / 3.14159265358979323846);
      amax = fmax(amax, ang);
      asum += (unsigned long long)(ang * 1e9 + 0.5);
      nv++;
      over1 += ang > 1.0 ? 1u : 0u;
      over10 += ang > 10.0 ? 1u : 0u;
    }
  }
  if (bad) atomicAdd(out + 0, bad);
  atomicMax(out + 1, (unsigned long long)__double_as_longlong(verr));
  atomicMax(out + 2, (unsigned long long)__double_as_longlong(amax));
  if (asum) atomicAdd(out + 3, asum);
  if (nv) atomicAdd(out + 4, nv);
  if (over1) atomicAdd(out + 5, over1);
  if (over10) atomicAdd(out + 6, over10);
}

// ----------------------------------------------------------------------
// rasterisation

// order-preserving integer images of the depths (larger depth = nearer)
__device__ __forceinline__ uint64_t ord64(double x) {
  if (x == 0.0) x = 0.0;
  const uint64_t u = __double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord64(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(u);
}
__device__ __forceinline__ uint32_t ord32(float x) {
  if (x == 0.0f) x = 0.0f;
  const uint32_t u = __float_as_uint(x);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord32(uint32_t k) {
  const uint32_t u = (k >> 31) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

// loads triangle t, projects it and sets up its reference box; false when
// a vertex is rejected by the projection
template <typename R>
__device__ __forceinline__ bool setup_tri(const MeshView<R>& m, uint64_t t, const ViewDev<R>& v, Projected<R>& P,
                                          R& area) {
#pragma unroll
  for (int n = 0; n < 3; n++) {
    const V3<R> p = {m.pos[t].p[3 * n + 0], m.pos[t].p[3 * n + 1], m.pos[t].p[3 * n + 2]};
    if (!project(v, p, P.q[n])) return false;
  }
  area = raster_box(v, P);
  return true;
}

// the image owning raster work item u (the last job starting at or before
// it: images with no triangles share their successor's start)
template <typename R>
__device__ __forceinline__ int find_rjob(const RasterJob<R>* __restrict__ jobs, int n, uint64_t u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].wk0 <= u) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ----------------------------------------------------------------------
// raster
//
// One triangle per thread; a workgroup round takes 256 consecutive work
// items of one image, so the image's constants are uniform.  Workgroups
// stride over the rounds and load the next round's triangles while drawing
// the current ones.  In FP64 the coverage loop runs over the narrowed box
// of tight_box.
// Modes:
//   kRasterFp32     FP32: packed (depth << 32 | id) keys, one pass
//   kRasterDepth64  FP64 depth pass that keeps the fragments that may end
//                   up nearest: a key below the value read from the z-buffer
//                   can never win (keys only grow); every other one goes
//                   through atomicMax and into a workgroup list in LDS, which
//                   is flushed with one global reservation per round.  The
//                   list holds every fragment whose key equals the final
//                   value.  Overflow (LDS or global) sets bit 2 of
//                   hdr->overflow.
//   kRasterId64     FP64 id pass over the geometry: the latest id (serial,
//                   or tri_id64 in the fused frames) at each pixel's final
//                   depth; runs only when the fragment list overflowed
constexpr int kFragLds = 1024;

// Appends fr to the workgroup's LDS fragment list (s_n entries, cap_lds
// slots), one LDS atomic per wave; fragments that do not fit go straight to
// the frame's list (gfrags, gcap entries, counted in hdr->nfrag).  Called
// by any subset of a wave's lanes.
template <typename FT>
__device__ __forceinline__ void append_frag(const FT& fr, FT* s_frag, uint32_t* s_n, uint32_t cap_lds,
                                            FT* __restrict__ gfrags, uint32_t gcap, FrameHdr* __restrict__ hdr) {
  const uint64_t act = __ballot(1);
  const int lane = int(__lane_id());
  const int leader = __ffsll((long long)act) - 1;
  uint32_t o0 = 0;
  if (lane == leader) o0 = atomicAdd(s_n, uint32_t(__popcll(act)));
  const uint32_t o = uint32_t(__shfl(int(o0), leader)) + uint32_t(__popcll(act & ((1ull << lane) - 1ull)));
  if (o < cap_lds) {
    s_frag[o] = fr;
  } else {
    const uint32_t g = atomicAdd(&hdr->nfrag, 1u);
    if (g < gcap) gfrags[g] = fr;
    else atomicOr(&hdr->overflow, 2u);
  }
}

// Projection (project.c), the reference's box and area (render.c:60-89)
// and the narrowed box of one triangle; returns its candidate pixels (0
// when a vertex fails to project or, in FP32, the projection has no area:
// such a triangle's weights are all +-0 and its depth 1 / +-0).
template <typename R>
__device__ __forceinline__ uint32_t setup_triangle(const ViewDev<R>& v, const V3<R> wpos[3], Projected<R>& P) {
  bool vis = true;
#pragma unroll
  for (int n = 0; n < 3; n++) vis = vis && project(v, wpos[n], P.q[n]);
  if (!vis) return 0;
  const R area = raster_box(v, P);
  if (sizeof(R) == 4 && area == R(0)) return 0;
  // about 12 reference-box candidates of a pixel-sized triangle narrow to 3
  tight_box(v, area, P);
  if (P.i1 < P.i0 || P.j1 < P.j0) return 0;
  // a box lies inside one image: fewer than 2^31 pixels
  return uint32_t(P.i1 - P.i0 + 1) * uint32_t(P.j1 - P.j0 + 1);
}

// The id type of a raster pass (uint32_t serial ids, or uint64_t
// tri_id64 in the fused FP64 frames) and its fragment record.
template <typename TId>
using FragOf = typename std::conditional<sizeof(TId) == 8, Frag64, Frag>::type;

__device__ __forceinline__ void atomic_max_id(uint32_t* p, uint32_t v) { atomicMax(p, v); }
__device__ __forceinline__ void atomic_max_id(uint64_t* p, uint64_t v) {
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// One candidate pixel (i, j) of triangle P (id tid1): the reference's
// barycentric test and harmonic depth (render.c:91-124), then the z-buffer
// update of `MODE`.  Depth64 and Append put fragments in the workgroup list
// s_frag (s_n entries, `cap_lds` slots); Append sends those that do not fit
// straight to the frame's list (gfrags, gcap entries, counted in
// hdr->nfrag).
template <typename R, int MODE, typename TId>
__device__ __forceinline__ void pixel_update(const ViewDev<R>& v, uint64_t row0, const Projected<R>& P, int32_t i,
                                             int32_t j, TId tid1, uint64_t* __restrict__ key, TId* __restrict__ id,
                                             FragOf<TId>* s_frag, uint32_t* s_n, uint32_t cap_lds,
                                             FragOf<TId>* __restrict__ gfrags, uint32_t gcap,
                                             FrameHdr* __restrict__ hdr) {
  R w0, w1, w2;
  if (!pixel_weights(P, pixel_centre<R>(i, v.fw), pixel_centre<R>(j, v.fh), w0, w1, w2)) return;
  const R dep = pixel_depth(P, w0, w1, w2);
  const uint64_t px = row0 + uint64_t(j) * uint64_t(v.W) + uint64_t(i);
  if constexpr (MODE == kRasterFp32) {
    // keys only grow: a key at or below a value read from the buffer can
    // never change it and skips the atomic
    const uint64_t k = uint64_t(ord32(float(dep))) << 32 | uint64_t(tid1);
    if (k > key[px]) atomicMax(reinterpret_cast<unsigned long long*>(key + px), (unsigned long long)k);
  } else if constexpr (MODE == kRasterAppend) {
    FragOf<TId> fr;
    fr.px = uint32_t(px);
    fr.id = tid1;
    fr.key = sizeof(R) == 8 ? ord64(double(dep)) : (uint64_t(ord32(float(dep))) << 32 | uint64_t(tid1));
    append_frag(fr, s_frag, s_n, cap_lds, gfrags, gcap, hdr);
  } else if constexpr (MODE == kRasterKey64) {
    const uint64_t k = ord64(double(dep));
    if (k > key[px]) atomicMax(reinterpret_cast<unsigned long long*>(key + px), (unsigned long long)k);
  } else if constexpr (MODE == kRasterDepth64) {
    const uint64_t k = ord64(double(dep));
    if (k < key[px]) return;
    atomicMax(reinterpret_cast<unsigned long long*>(key + px), (unsigned long long)k);
    // every fragment whose key equals the final value is kept: its key was
    // at least the value read before its own atomic
    const uint32_t o = atomicAdd(s_n, 1u);
    if (o < cap_lds) {
      FragOf<TId> fr;
      fr.px = uint32_t(px);
      fr.id = tid1;
      fr.key = k;
      s_frag[o] = fr;
    }
  } else {
    if (ord64(double(dep)) == key[px]) atomic_max_id(id + px, tid1);
  }
}

// one triangle drawn by its own lane: every candidate of its box in turn
template <typename R, int MODE, typename TId>
__device__ __forceinline__ void draw_triangle(const ViewDev<R>& v, uint64_t row0, const V3<R> wpos[3], TId tid1,
                                              uint64_t* __restrict__ key, TId* __restrict__ id,
                                              FragOf<TId>* s_frag, uint32_t* s_n, uint32_t cap_lds,
                                              FragOf<TId>* __restrict__ gfrags = nullptr, uint32_t gcap = 0,
                                              FrameHdr* __restrict__ hdr = nullptr) {
  Projected<R> P;
  if (!setup_triangle<R>(v, wpos, P)) return;
  for (int32_t j = P.j0; j <= P.j1; j++)
    for (int32_t i = P.i0; i <= P.i1; i++)
      pixel_update<R, MODE, TId>(v, row0, P, i, j, tid1, key, id, s_frag, s_n, cap_lds, gfrags, gcap, hdr);
}

// registers of one raster work item, loaded a round ahead
template <typename R>
struct TriRegs {
  R p[9];
  uint32_t ed;
  uint64_t t;
  int jb;
  bool ok;
};

template <typename R>
__device__ __forceinline__ void fetch_tri(const MeshView<R>& m, const RasterJob<R>* __restrict__ rjobs, int njobs,
                                          uint64_t u0, TriRegs<R>& T) {
  T.jb = njobs > 1 ? find_rjob(rjobs, njobs, u0) : 0;
  T.t = rjobs[T.jb].tri0 + (u0 + threadIdx.x - rjobs[T.jb].wk0);
  T.ok = T.t < rjobs[T.jb].tri1;
  T.ed = 0;
  if (T.ok) {
    T.ed = edges_of(m, T.t);
#pragma unroll
    for (int n = 0; n < 9; n++) T.p[n] = m.pos[T.t].p[n];
  }
}

template <typename R, int MODE>
__global__ __launch_bounds__(kThreads) void k_raster(MeshView<R> m, const RasterJob<R>* __restrict__ rjobs, int njobs,
                                                     FrameHdr* __restrict__ hdr, uint64_t* __restrict__ key,
                                                     uint32_t* __restrict__ id, Frag* __restrict__ frags,
                                                     uint32_t cap) {
  __shared__ Frag s_frag[MODE == kRasterDepth64 ? kFragLds : 1];
  __shared__ uint32_t s_n, s_base;
  // a frame whose triangles outgrew the mesh (bit 1, set by the plan of a
  // replayed frame) has triangles that were never stored: draw nothing
  if (hdr->overflow & 1u) return;
  if (MODE == kRasterId64 && !(hdr->overflow & 2u)) return;
  const uint64_t nwork = hdr->nwork;
  const uint64_t stride = uint64_t(gridDim.x) * kThreads;
  uint64_t u0 = uint64_t(blockIdx.x) * kThreads;
  TriRegs<R> cur;
  cur.ok = false;
  if (u0 < nwork) fetch_tri(m, rjobs, njobs, u0, cur);
  for (; u0 < nwork; u0 += stride) {
    // the next round's triangle is loaded while this one is drawn
    TriRegs<R> nxt;
    nxt.ok = false;
    if (u0 + stride < nwork) fetch_tri(m, rjobs, njobs, u0 + stride, nxt);
    if (MODE == kRasterDepth64) {
      if (threadIdx.x == 0) s_n = 0;
      __syncthreads();
    }
    if (cur.ok && ((cur.ed >> 15) & 1u)) {
      const V3<R> wp[3] = {{cur.p[0], cur.p[1], cur.p[2]}, {cur.p[3], cur.p[4], cur.p[5]},
                           {cur.p[6], cur.p[7], cur.p[8]}};
      const ViewDev<R> v = rjobs[cur.jb].v;
      draw_triangle<R, MODE>(v, rjobs[cur.jb].px0, wp, uint32_t(cur.t + 1), key, id, s_frag, &s_n, kFragLds);
    }
    if constexpr (MODE == kRasterDepth64) {
      __syncthreads();
      const uint32_t n = s_n < uint32_t(kFragLds) ? s_n : uint32_t(kFragLds);
      if (threadIdx.x == 0) {
        s_base = n ? atomicAdd(&hdr->nfrag, n) : 0u;
        if (s_n > uint32_t(kFragLds) || (n && uint64_t(s_base) + n > cap)) atomicOr(&hdr->overflow, 2u);
      }
      __syncthreads();
      const uint32_t b0 = s_base;
      for (uint32_t e = threadIdx.x; e < n; e += kThreads) {
        if (uint64_t(b0) + e < cap) frags[b0 + e] = s_frag[e];
      }
      __syncthreads();
    }
    cur = nxt;
  }
}

// ----------------------------------------------------------------------
// frames without a stored mesh: tessellation fused with the raster
//
// Each tile is classified and scanned in chunks of 256 cells as in the emit
// pass, but a chunk holds only about a third as many triangles as cells, so
// its triangles are not drawn in place.  They are queued instead, as
// compact items (serial position in FP32, cell, vertex corners,
// threshold), in a per-workgroup LDS ring that carries them over chunk and
// tile boundaries (within one field pass), and the ring is drained 256
// items at a time:
// every lane then computes, projects and draws one triangle.  A drained
// triangle rereads the corner values its vertices need (they were loaded
// moments earlier by the same workgroup, so they come from cache).  An FP32
// frame stores only the cell index of each triangle (at its serial
// position), an FP64 frame nothing (its ids name the cell and slot); the
// resolve recomputes what it needs from the field.  Triangles are drawn
// into every image of the field pass that shows their object (jimg:
// offsets per field pass into a list of the images showing any of its
// objects).  In Append mode every covered pixel collects in an LDS list
// that is flushed to the frame's list once half full; the z-test modes run
// only after that list overflowed.  A frame flagged by the plan as outgrowing `cap`
// draws nothing (it is rendered again).
constexpr uint32_t kFragLdsFused = 512;  // LDS fragment slots of a workgroup (about 20 per drain at 512^3)
constexpr uint32_t kFragFlushAt = 256;   // flushed to the frame's list from this fill on
constexpr uint32_t kRing = 512;          // a power of two, at least 2 * kThreads

// a queued triangle names its threshold in 2 bits (ring item bits 18-19),
// and the drains index ThrSet::thr and ::obj with it
static_assert(kMaxThr == 4, "the ring's threshold field is 2 bits wide");

struct FusedLds {
  uint8_t code[kMaxThr][kThreads];
  uint16_t pref[kMaxThr][kThreads];
  uint16_t low[kTileChunks][kThreads];  // lower-plane codes of the tile's cells (cell p at [p / 256][p % 256])
  uint64_t wave[kThreads / 64];
  uint8_t ntri[256];
  uint32_t rg[kRing];  // serial position of the queued triangle
  uint32_t rc[kRing];  // its cell
  uint32_t re[kRing];  // its vertices' (tail, head) corners (slot entry bits 0-17) | threshold << 18 |
                       // slot << 20 | corner code << 24
  uint32_t nfrag, fbase, fcnt;
};

// positions of the triangle whose vertices join the corners in `ent` (slot
// entry bits 0-17) of cell L, reading the two corner values of each
// vertex's lattice edge from the field; on a grid lattice the cell's six
// coordinates are read once.  Every load depends only on L and ent, so they
// are in flight together.
template <typename R>
__device__ __forceinline__ void slot_positions_mem(const FieldDev& f, const CellLoc& L, double thr, uint32_t ent,
                                                   V3<R> P[3]) {
  const C3D_GLOBAL double* p = gptr(f.f) + L.p0;
  if (f.cx) {
#pragma unroll
    for (int vv = 0; vv < 3; vv++) {
      const uint32_t tail = (ent >> (6 * vv)) & 7u;
      const uint32_t head = (ent >> (6 * vv + 3)) & 7u;
      P[vv] = interp<R>(thr, corner_pos<R>(f, L, tail), p[corner_offset(f, tail)], corner_pos<R>(f, L, head),
                        p[corner_offset(f, head)]);
    }
    return;
  }
  const double x0 = gptr(f.gx)[L.i], x1 = gptr(f.gx)[L.i + 1], y0 = gptr(f.gy)[L.j], y1 = gptr(f.gy)[L.j + 1];
  const double z0 = gptr(f.gz)[L.k], z1 = gptr(f.gz)[L.k + 1];
#pragma unroll
  for (int vv = 0; vv < 3; vv++) {
    const uint32_t tail = (ent >> (6 * vv)) & 7u;
    const uint32_t head = (ent >> (6 * vv + 3)) & 7u;
    const V3<R> pt = {R(tail & 1u ? x1 : x0), R(tail & 2u ? y1 : y0), R(tail & 4u ? z1 : z0)};
    const V3<R> ph = {R(head & 1u ? x1 : x0), R(head & 2u ? y1 : y0), R(head & 4u ? z1 : z0)};
    P[vv] = interp<R>(thr, pt, p[corner_offset(f, tail)], ph, p[corner_offset(f, head)]);
  }
}

// Ends a drain: in Append mode, moves the workgroup's LDS fragments to the
// frame's list once they fill kFragFlushAt slots (every time with `all`).
// One reservation on the frame's counter per flush keeps that shared atomic
// rare.
template <int MODE, typename FT>
__device__ __forceinline__ void flush_frags(FusedLds& S, FT* s_frag, FrameHdr* __restrict__ hdr,
                                            FT* __restrict__ frags, uint32_t fcap, bool all) {
  __syncthreads();
  if constexpr (MODE == kRasterAppend) {
    if (threadIdx.x == 0) {
      const uint32_t have = S.nfrag;
      const uint32_t c = have < kFragLdsFused ? have : kFragLdsFused;
      const bool go = c && (all || c >= kFragFlushAt);
      uint32_t b = 0;
      if (go) {
        b = atomicAdd(&hdr->nfrag, c);
        if (uint64_t(b) + c > fcap) atomicOr(&hdr->overflow, 2u);
        S.nfrag = 0;
      }
      S.fbase = b;
      S.fcnt = go ? c : 0u;
    }
    __syncthreads();
    const uint32_t b0 = S.fbase, c = S.fcnt;
    if (c) {
      for (uint32_t e = threadIdx.x; e < c; e += kThreads) {
        if (uint64_t(b0) + e < fcap) frags[b0 + e] = s_frag[e];
      }
      __syncthreads();
    }
  }
}

// Draws the n (<= kThreads) oldest queued triangles of field pass jb, one
// per thread: with 64-bit ids (FP64) as tri_id64, else as their serial
// position + 1.
template <typename R, int MODE, typename TId>
__device__ __forceinline__ void drain_ring(FusedLds& S, FragOf<TId>* s_frag, uint32_t head, uint32_t n,
                                           const TessJob* __restrict__ jobs, uint32_t jb,
                                           const RasterJob<R>* __restrict__ rjobs, const uint32_t* __restrict__ jimg,
                                           FrameHdr* __restrict__ hdr, uint64_t* __restrict__ key,
                                           TId* __restrict__ id, FragOf<TId>* __restrict__ frags, uint32_t fcap) {
  V3<R> P[3];
  bool drawn = false;
  uint32_t obj = 0;
  TId tid = 0;
  if (threadIdx.x < n) {
    const FieldDev f = jobs[jb].f;
    const uint32_t s = (head + threadIdx.x) & (kRing - 1u);
    const uint32_t cc = S.rc[s], x = S.re[s];
    if constexpr (sizeof(TId) == 4) tid = S.rg[s] + 1u;
    const CellLoc L = locate(f, cc);
    if (interior_cell(f, L)) {
      const uint32_t q = (x >> 18) & 3u;
      // one load per lane indexed by the threshold number: fewer vector
      // instructions and spilled registers than selecting among kMaxThr
      const double thr = jobs[jb].ts.thr[q];
      obj = jobs[jb].ts.obj[q];
      slot_positions_mem<R>(f, L, thr, x & 0x3ffffu, P);
      if constexpr (sizeof(TId) == 8) tid = tri_id64(obj, cc, (x >> 20) & 15u);
      drawn = true;
    }
  }
  // the images of this field pass (the same for the whole workgroup)
  for (uint32_t u = jimg[jb]; u < jimg[jb + 1]; u++) {
    const uint32_t k = jimg[u];
    const bool here = drawn && obj >= rjobs[k].obj0 && obj < rjobs[k].obj1;
    const ViewDev<R> v = rjobs[k].v;
    if (here)
      draw_triangle<R, MODE, TId>(v, rjobs[k].px0, P, tid, key, id, s_frag, &S.nfrag, kFragLdsFused, frags, fcap,
                                  hdr);
  }
  flush_frags<MODE>(S, s_frag, hdr, frags, fcap, false);
}

// The work-item loop shared by the queue kernels: marches each work item's
// column up its planes, classifies each tile's cells in chunks of 256
// (loading only the upper corners: the lower-plane codes carry over from the
// plane below, in LDS), queues the chunk's triangles in the LDS ring in
// serial order and calls drain(head, n, jb) for every full 256 items (and
// for the rest when the field pass changes and at the end).  With kPlanes
// each plane's upper codes are first computed for the whole tile by column
// segments (classify_plane, into s_hi: the cells of a segment share their
// rows of lattice points) and the chunks read them from LDS; the fused
// kernel measured 3.6% (FP64) and 1.3% (FP32) faster so, the FP32 emit
// 1.5%; the FP64 emit, which would lose a workgroup per CU to the extra
// LDS, loads each cell's corners itself.  With
// kStoreCell the cell index of each triangle is stored at its serial
// position as it is queued; with kStoreCode every cell's corner code of
// every threshold q is stored at codeout[code_off + q * ncells + cell]: by
// each cell's thread as bytes (kStoreCode 1) or from LDS as 4-byte words
// (kStoreCode 2; measured faster in the FP64 emit, slower in the FP32 one).
// With kSerial each queued triangle carries its serial position (from the
// tile starts `start`); without it (fused FP64 frames, whose ids come from
// the cell and slot) start is not read.  Expects S.ntri loaded and a
// barrier since.
template <bool kStoreCell, int kStoreCode, bool kReload, bool kPlanes, bool kSerial, typename Drain>
__device__ __forceinline__ void ring_tiles(FusedLds& S, uint16_t* s_hi, const TessJob* __restrict__ jobs, int njobs,
                                           uint32_t nitems, const uint64_t* __restrict__ start,
                                           uint32_t* __restrict__ cellout, uint64_t cap, uint8_t* __restrict__ codeout,
                                           Drain&& drain) {
  // ring state, the same in every thread: oldest item, items queued, and
  // the field pass they belong to
  uint32_t head = 0, npend = 0, cur = 0;
  const XcdWalk W = xcd_walk(jobs, njobs, nitems);
  for (uint32_t pp = W.begin; pp < W.end; pp += W.step) {
    uint32_t jb, item;
    if (!xcd_item(jobs, njobs, nitems, W.x, pp, jb, item)) continue;
    // the queue holds one field pass at a time
    while (npend && jb != cur) {
      const uint32_t n = npend < kThreads ? npend : kThreads;
      drain(head, n, cur);
      head = (head + n) & (kRing - 1u);
      npend -= n;
    }
    cur = jb;
    // The pass's field and thresholds are read through jbx, which with
    // kReload is made opaque after every drain: the compiler then reloads
    // them (scalar loads from the constant cache) instead of holding them in
    // scalar registers across the drain, where the fused kernel's image
    // constants need those registers (it had spilled scalars to vector lanes
    // at twice the rate; the frames got 2-3.5% faster, the emit, whose
    // drains hold no image constants, 3% slower).
    uint32_t jbx = jb;
    const auto Fd = [&]() -> const FieldDev& { return jobs[jbx].f; };
    const auto Ts = [&]() -> const ThrSet& { return jobs[jbx].ts; };
    uint32_t col, k0, k1;
    item_range(Fd(), item, col, k0, k1);
    // the tile's span and plane strides, recomputed from the pass (and so
    // after a drain) rather than held
    const auto Tsp = [&]() { return tile_span(Fd(), k0, col); };
    const uint32_t nch = (Tsp().n + kThreads - 1) / kThreads;
    const auto nxy = [&]() { return uint64_t(Fd().nx) * Fd().ny; };
    const auto ncxy = [&]() { return Fd().ncx * Fd().ncy; };
    // codes of the first plane's corners: with kPlanes by column segments
    // (the cells of a segment share their rows of lattice points), else
    // each thread its own cells (reading back only what it wrote)
    if constexpr (kPlanes) {
      __syncthreads();  // the previous item's chunk loop is done with S.low
      classify_plane(Fd(), Ts(), Tsp(), gptr(Fd().f) + Tsp().l0, &S.low[0][0]);
    } else {
      for (uint32_t ch = 0; ch < nch; ch++) {
        const uint32_t p = ch * kThreads + threadIdx.x;
        if (p < Tsp().n) {
          double v[4];
          load_plane4(gptr(Fd().f) + Tsp().l0 + tile_corner_off(Fd(), Tsp(), p), Fd().nx, v);
          S.low[ch][threadIdx.x] = uint16_t(plane_codes(Ts(), v));
        }
      }
    }
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t tile = k * Fd().tg.cols + col;
      const uint32_t ctile = Tsp().c0 + (k - k0) * ncxy();
      const C3D_GLOBAL double* up = gptr(Fd().f) + Tsp().l0 + uint64_t(k - k0 + 1) * nxy();
      if constexpr (kPlanes) {
        __syncthreads();  // the previous plane's chunk loop is done with s_hi
        classify_plane(Fd(), Ts(), Tsp(), up, s_hi);
        __syncthreads();
      }
      uint64_t base[kMaxThr];
#pragma unroll
      for (int q = 0; q < kMaxThr; q++) base[q] = kSerial && q < Ts().n ? start[Ts().boff[q] + tile] : 0;
      for (uint32_t ch = 0; ch < nch; ch++) {
        const uint32_t c0 = ctile + ch * kThreads;
        const uint32_t p = ch * kThreads + threadIdx.x;
        const bool valid = p < Tsp().n;
        uint32_t code[kMaxThr];
        uint64_t packed = 0;
#pragma unroll
        for (int q = 0; q < kMaxThr; q++) code[q] = 0;
        if (valid) {
          uint32_t hi;
          if constexpr (kPlanes) {
            hi = s_hi[p];
          } else {
            double v[4];
            load_plane4(up + tile_corner_off(Fd(), Tsp(), p), Fd().nx, v);
            hi = plane_codes(Ts(), v);
          }
          const uint32_t lo = S.low[ch][threadIdx.x];
          S.low[ch][threadIdx.x] = uint16_t(hi);
#pragma unroll
          for (int q = 0; q < kMaxThr; q++) {
            const uint32_t cd = ((lo >> (4 * q)) & 15u) | ((hi >> (4 * q)) & 15u) << 4;
            code[q] = q < Ts().n ? cd : 0u;
            packed |= uint64_t(S.ntri[code[q]]) << (16 * q);
          }
          if (kStoreCode == 1) {
            uint8_t* co = codeout + jobs[jb].code_off + (c0 + threadIdx.x);
#pragma unroll
            for (int q = 0; q < kMaxThr; q++)
              if (q < Ts().n) co[uint64_t(q) * Fd().ncells] = uint8_t(code[q]);
          }
        }
        uint64_t tot;
        const uint64_t ex = block_excl_scan64(packed, S.wave, tot);
        uint32_t qoff[kMaxThr + 1];
        qoff[0] = 0;
#pragma unroll
        for (int q = 0; q < kMaxThr; q++) {
          qoff[q + 1] = qoff[q] + (uint32_t(tot >> (16 * q)) & 0xffffu);
          S.pref[q][threadIdx.x] = uint16_t(ex >> (16 * q));
          if (valid) S.code[q][threadIdx.x] = uint8_t(code[q]);
        }
        __syncthreads();
        if (kStoreCode == 2) {
          // the chunk's codes of threshold q are bytes [d0, d0 + nv) of the
          // code array: whole 4-byte words inside that range are stored as
          // words, the bytes of the (at most two) words it shares with the
          // neighbouring chunks one by one
          const uint32_t nv = Tsp().n - ch * kThreads < kThreads ? Tsp().n - ch * kThreads : kThreads;
          const uint32_t nw = (kThreads + 3u) / 4u + 1u;  // words touched by 256 bytes at any alignment
          for (uint32_t w = threadIdx.x; w < uint32_t(Ts().n) * nw; w += kThreads) {
            const uint32_t q = w / nw, kw = w - q * nw;
            const uint64_t d0 = jobs[jb].code_off + uint64_t(q) * Fd().ncells + c0;
            const uint64_t a = (d0 & ~uint64_t(3)) + 4u * kw;
            const uint64_t lo = a > d0 ? a : d0, hi = a + 4u < d0 + nv ? a + 4u : d0 + nv;
            if (lo >= hi) continue;
            if (lo == a && hi == a + 4u) {
              const uint32_t l = uint32_t(a - d0);
              const uint32_t word = uint32_t(S.code[q][l]) | uint32_t(S.code[q][l + 1]) << 8 |
                                    uint32_t(S.code[q][l + 2]) << 16 | uint32_t(S.code[q][l + 3]) << 24;
              *reinterpret_cast<uint32_t*>(codeout + a) = word;
            } else {
              for (uint64_t b = lo; b < hi; b++) codeout[b] = S.code[q][uint32_t(b - d0)];
            }
          }
        }
        const uint32_t total = qoff[kMaxThr];
        for (uint32_t pos = 0; pos < total;) {
          // queue as many of the chunk's triangles as the ring holds, in
          // serial order (threshold, cell, slot)
          const uint32_t room = kRing - npend;
          const uint32_t take = total - pos < room ? total - pos : room;
          for (uint32_t e = threadIdx.x; e < take; e += kThreads) {
            const uint32_t it = pos + e;
            int q = 0;
#pragma unroll
            for (int z = 1; z < kMaxThr; z++) q += it >= qoff[z] ? 1 : 0;
            uint32_t qsel = 0;
#pragma unroll
            for (int z = 1; z < kMaxThr; z++) qsel = q == z ? qoff[z] : qsel;
            const uint32_t r = it - qsel;
            uint32_t l = 0;
#pragma unroll
            for (uint32_t step = kThreads / 2; step; step >>= 1) l += S.pref[q][l + step] <= r ? step : 0u;
            const uint32_t sl = r - S.pref[q][l];
            const uint32_t cc = c0 + l;
            const uint32_t cd = S.code[q][l];
            const uint32_t s = (head + npend + e) & (kRing - 1u);
            if constexpr (kSerial) {
              uint64_t bq = base[0];
#pragma unroll
              for (int z = 1; z < kMaxThr; z++) bq = q == z ? base[z] : bq;
              const uint64_t g = bq + r;
              if (kStoreCell && g < cap) cellout[g] = cc;
              S.rg[s] = uint32_t(g);
            }
            S.rc[s] = cc;
            S.re[s] = (uint32_t(c_cslot[cd * 12u + sl]) & 0x3ffffu) | uint32_t(q) << 18 | sl << 20 | cd << 24;
          }
          pos += take;
          npend += take;
          __syncthreads();
          while (npend >= kThreads) {
            drain(head, kThreads, jb);
            if constexpr (kReload) asm volatile("" : "+s"(jbx));
            head = (head + kThreads) & (kRing - 1u);
            npend -= kThreads;
          }
        }
        // the next chunk's triangles follow this chunk's in every object
#pragma unroll
        for (int q = 0; q < kMaxThr; q++) base[q] += qoff[q + 1] - qoff[q];
      }
    }
  }
  while (npend) {
    const uint32_t n = npend < kThreads ? npend : kThreads;
    drain(head, n, cur);
    head = (head + n) & (kRing - 1u);
    npend -= n;
  }
}

// Occupancy: the FP64 form runs five waves per SIMD (its LDS allows no
// more), the FP32 form seven (72 VGPRs, no spill; 3% faster than six).
template <typename R, int MODE>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 || MODE == kRasterAppend ? 5 : 7))) void
k_tess_raster(const TessJob* __restrict__ jobs, int njobs, uint32_t nitems, const uint64_t* __restrict__ start,
              uint32_t* __restrict__ cellout, uint64_t cap, const RasterJob<R>* __restrict__ rjobs,
              const uint32_t* __restrict__ jimg, FrameHdr* __restrict__ hdr, uint64_t* __restrict__ key,
              void* __restrict__ idv, void* __restrict__ fragv, uint32_t fcap) {
  static_assert(MODE != kRasterDepth64, "the fused kernel keeps every fragment (kRasterAppend)");
  // FP64 frames name triangles by tri_id64, FP32 ones by serial position
  using TId = typename std::conditional<sizeof(R) == 8, uint64_t, uint32_t>::type;
  constexpr bool kSerial = sizeof(R) == 4;
  TId* id = static_cast<TId*>(idv);
  FragOf<TId>* frags = static_cast<FragOf<TId>*>(fragv);
  constexpr bool kPlanes = true;
  __shared__ FusedLds S;
  __shared__ uint16_t s_hi[kPlanes ? kTileCells : 1];
  __shared__ FragOf<TId> s_frag[MODE == kRasterAppend ? kFragLdsFused : 1];
  if (hdr->overflow & 1u) return;
  // the FP64 z-test passes run only after the fragment list overflowed
  if ((MODE == kRasterKey64 || MODE == kRasterId64) && !(hdr->overflow & 2u)) return;
  S.ntri[threadIdx.x] = c_ntri[threadIdx.x];
  if (threadIdx.x == 0) S.nfrag = 0;
  __syncthreads();
  uint64_t ndrawn = 0;  // triangles drained (the same in every thread)
  // the first pass of an FP32 frame stores the cell indices
  ring_tiles<kSerial && (MODE == kRasterAppend || MODE == kRasterFp32), 0, true, kPlanes, kSerial>(
      S, s_hi, jobs, njobs, nitems, start, cellout, cap, nullptr, [&](uint32_t head, uint32_t n, uint32_t jb) {
        ndrawn += n;
        drain_ring<R, MODE, TId>(S, s_frag, head, n, jobs, jb, rjobs, jimg, hdr, key, id, frags, fcap);
      });
  flush_frags<MODE>(S, s_frag, hdr, frags, fcap, true);
  // an FP64 frame's triangle count, which no count pass gives it
  if (!kSerial && MODE == kRasterAppend && threadIdx.x == 0 && ndrawn)
    atomicAdd(reinterpret_cast<unsigned long long*>(&hdr->ntri), (unsigned long long)ndrawn);
}

// ----------------------------------------------------------------------
// tessellation: emit pass (the stored mesh)
//
// The work-item loop of the queue kernels; every drain computes up to 256
// triangles, one per thread, into records in LDS (a position record and a
// gather record each), and the workgroup then stores those records at
// their serial positions as 8-byte words, thread q taking words q, q + 256,
// ... of each array: the drain's triangles are a few runs of consecutive
// positions (one per object in the drained chunks), so every store
// instruction writes contiguous bytes and whole lines, where one store per
// record field (14 streams) had written scattered 4- or 8-byte pieces.  The
// stores are non-temporal: the mesh streams past the caches, and the field
// planes the workgroups march through stay in L2.
//
// Every earlier vector memory operation is waited for once before the
// drain's stores, so that nothing is in flight when they issue: the
// compiler then issues the store sequence back to back, where otherwise it
// waited between stores for loads still counted ahead of them.

// stores n records of type Rec from LDS at the drain's serial positions
template <typename Rec>
__device__ __forceinline__ void store_drain(const FusedLds& S, const Rec* s_src, uint32_t head, uint32_t n,
                                            Rec* out, uint64_t cap) {
  constexpr uint32_t kWords = sizeof(Rec) / 8u;
  const uint64_t* src = reinterpret_cast<const uint64_t*>(s_src);
  char* dst = reinterpret_cast<char*>(out);
  for (uint32_t k = threadIdx.x; k < n * kWords; k += kThreads) {
    const uint32_t r = k / kWords;
    const uint32_t g = S.rg[(head + r) & (kRing - 1u)];
    if (g < cap)
      __builtin_nontemporal_store(src[k], reinterpret_cast<uint64_t*>(dst + uint64_t(g) * sizeof(Rec) +
                                                                     8u * (k - r * kWords)));
  }
}

template <typename R>
__device__ __forceinline__ void emit_ring(FusedLds& S, MeshPos<R>* s_pos, MeshAux<R>* s_aux, uint32_t head,
                                          uint32_t n, const TessJob* __restrict__ jobs, uint32_t jb,
                                          const MeshView<R>& m, FrameHdr* __restrict__ hdr) {
  if (threadIdx.x < n) {
    const FieldDev f = jobs[jb].f;
    const uint32_t s = (head + threadIdx.x) & (kRing - 1u);
    const uint32_t cc = S.rc[s], x = S.re[s];
    const CellLoc L = locate(f, cc);
    const uint32_t q = (x >> 18) & 3u;
    const double thr = jobs[jb].ts.thr[q];
    V3<R> P[3];
    slot_positions_mem<R>(f, L, thr, x & 0x3ffffu, P);
    const V3<R> w = smoothing_weight<R>(P);
    const uint32_t ent_edges = uint32_t(c_cslot[(x >> 24) * 12u + ((x >> 20) & 15u)] >> 18) & 0x7fffu;
    const uint32_t edges = ent_edges | uint32_t(interior_cell(f, L)) << 15;
    MeshPos<R>& rp = s_pos[threadIdx.x];
    MeshAux<R>& ra = s_aux[threadIdx.x];
#pragma unroll
    for (int vv = 0; vv < 3; vv++) {
      rp.p[3 * vv + 0] = P[vv].x;
      rp.p[3 * vv + 1] = P[vv].y;
      rp.p[3 * vv + 2] = P[vv].z;
    }
    ra.w[0] = w.x;
    ra.w[1] = w.y;
    ra.w[2] = w.z;
    ra.cell = cc;
    if constexpr (sizeof(R) == 8) ra.edges = edges;
    else rp.edges = edges;
    if (uint64_t(S.rg[s]) >= m.cap) hdr->overflow = 1u;  // capacity reserved ahead of a replayed frame was too small
  }
  __syncthreads();
  __builtin_amdgcn_s_waitcnt(0);
  store_drain(S, s_pos, head, n, m.pos, m.cap);
  store_drain(S, s_aux, head, n, m.aux, m.cap);
  __syncthreads();
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_emit(const TessJob* __restrict__ jobs, int njobs, uint32_t nitems,
                                                   const uint64_t* __restrict__ start, MeshView<R> m,
                                                   FrameHdr* __restrict__ hdr) {
  __shared__ FusedLds S;
  __shared__ MeshPos<R> s_pos[kThreads];
  __shared__ MeshAux<R> s_aux[kThreads];
  constexpr bool kPl = sizeof(R) == 4;  // the FP64 build would lose a workgroup per CU to the LDS
  __shared__ uint16_t s_hi[kPl ? kTileCells : 1];
  S.ntri[threadIdx.x] = c_ntri[threadIdx.x];
  __syncthreads();
  ring_tiles<false, sizeof(R) == 8 ? 2 : 1, false, kPl, true>(S, s_hi, jobs, njobs, nitems, start, nullptr, 0, m.code,
                                                          [&](uint32_t head, uint32_t n, uint32_t jb) {
    emit_ring<R>(S, s_pos, s_aux, head, n, jobs, jb, m, hdr);
  });
}

// the latest triangle among the kept fragments at each pixel's final depth
template <typename TId>
__global__ void k_frag_ids(const FragOf<TId>* __restrict__ frags, uint32_t cap, const FrameHdr* __restrict__ hdr,
                           const uint64_t* __restrict__ key, TId* __restrict__ id) {
  if (hdr->overflow & 2u) return;
  const uint32_t n = min(hdr->nfrag, cap);
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    const FragOf<TId> fr = frags[f];
    if (fr.key == key[fr.px]) atomic_max_id(id + fr.px, fr.id);
  }
}

// the z-buffer keys of every listed fragment
template <typename FT>
__global__ void k_frag_keys(const FT* __restrict__ frags, uint32_t cap, const FrameHdr* __restrict__ hdr,
                            uint64_t* __restrict__ key) {
  if (hdr->overflow & 2u) return;
  const uint32_t n = min(hdr->nfrag, cap);
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    const uint64_t fkey = frags[f].key;
    const uint32_t fpx = frags[f].px;
    // keys only grow, so a value read here (however stale) that is at least
    // this key means the atomic could not change the pixel: skip it
    if (fkey > key[fpx]) atomicMax(reinterpret_cast<unsigned long long*>(key + fpx), (unsigned long long)fkey);
  }
}

// also clears the frame header's raster state (fragment count, list
// overflow) and, fresh, its triangle count and capacity bit
template <typename TId>
__global__ void k_zinit(uint64_t* __restrict__ key, TId* __restrict__ id, uint64_t npx, uint64_t k0, bool fresh,
                        FrameHdr* __restrict__ hdr) {
  if (hdr && blockIdx.x == 0 && threadIdx.x == 0) {
    hdr->nfrag = 0;
    hdr->overflow &= fresh ? 0u : 1u;
    if (fresh) hdr->ntri = 0;
  }
  for (uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < npx; p += uint64_t(gridDim.x) * blockDim.x) {
    key[p] = k0;
    if (id) id[p] = 0;
  }
}

// Line fragments (pixels and depths computed on the host, in drawing
// order) take part in the same z-test; their ids follow every triangle's
// (64-bit: after every object, kLineObj; 32-bit: after the frame's ntri).
template <typename TId>
__global__ void k_line_frags(const uint32_t* __restrict__ pix, const double* __restrict__ depth,
                             const FrameHdr* __restrict__ hdr, int pass, int fp32, uint64_t* __restrict__ key,
                             TId* __restrict__ id) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= hdr->nline) return;
  const uint32_t px = pix[f];
  const TId lid = sizeof(TId) == 8 ? TId(kLineObj << 40 | f) : TId(hdr->ntri + 1 + f);
  if (fp32) {
    const uint64_t k = uint64_t(ord32(float(depth[f]))) << 32 | uint64_t(lid);
    atomicMax(reinterpret_cast<unsigned long long*>(key + px), (unsigned long long)k);
    return;
  }
  const uint64_t k = ord64(depth[f]);
  if (pass == 0) {
    atomicMax(reinterpret_cast<unsigned long long*>(key + px), (unsigned long long)k);
  } else if (k == __hip_atomic_load(key + px, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    atomic_max_id(id + px, lid);
  }
}

// Final colour of every pixel from the winning fragment.  The vertex normals
// of a winning triangle are gathered here, on demand: the image needs them
// only where a triangle is nearest, so the smoothing of normal.c is done for
// those triangles alone, with the same sums.  rgb rows are written top-down
// (y index H-1 first) as the PPM writer packs them (output.c:113-123).
template <typename R>
__global__ __launch_bounds__(kThreads) void k_resolve(MeshView<R> m, const RasterJob<R>* __restrict__ rjobs, int njobs,
                                                      uint64_t npx_total, const ObjMeta* __restrict__ meta, int nobj,
                                                      const uint8_t* __restrict__ frag_rgb,
                                                      const FrameHdr* __restrict__ hdr,
                                                      const uint64_t* __restrict__ key,
                                                      const uint32_t* __restrict__ idbuf) {
  // a frame whose triangles outgrew the mesh is rendered again: skip it
  if (hdr->overflow & 1u) return;
  const uint64_t ntri = hdr->ntri;
  for (uint64_t p0 = uint64_t(blockIdx.x) * kThreads; p0 < npx_total; p0 += uint64_t(gridDim.x) * kThreads) {
    // images start on workgroup boundaries of the pixel space, so the
    // image of a workgroup round is uniform
    int jb = 0;
    if (njobs > 1) {
      int lo = 0, hi = njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rjobs[mid].px0 <= p0) lo = mid;
        else hi = mid - 1;
      }
      jb = lo;
    }
    const RasterJob<R>& J = rjobs[jb];
    const ViewDev<R> v = J.v;
    const uint64_t p = p0 + threadIdx.x;
    const uint64_t lp = p - J.px0;
    if (lp >= uint64_t(v.W) * uint64_t(v.H)) continue;
    const int32_t j = int32_t(lp / uint64_t(v.W));
    const int32_t i = int32_t(lp - uint64_t(j) * uint64_t(v.W));
    const uint64_t k = key[p];
    uint32_t id;
    double dep;
    if constexpr (sizeof(R) == 8) {
      id = idbuf[p];
      dep = unord64(k);
    } else {
      id = uint32_t(k);
      dep = id ? double(unord32(uint32_t(k >> 32))) : -DBL_MAX;
    }
    uint8_t r = J.bg[0], gg = J.bg[1], b = J.bg[2];
    if (id != 0 && id <= ntri) {
      const uint64_t t = id - 1;
      int lo = 0, hi = nobj - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (meta[mid].base <= t) lo = mid;
        else hi = mid - 1;
      }
      const ObjMeta om = meta[lo];
      Projected<R> P;
      R area;
      setup_tri(m, t, v, P, area);
      R w0, w1, w2;
      pixel_weights(P, pixel_centre<R>(i, v.fw), pixel_centre<R>(j, v.fh), w0, w1, w2);
      const uint32_t c = m.aux[t].cell;
      const uint32_t ed = edges_of(m, t);
      V3<R> nv[3];
#pragma unroll 1
      for (int n = 0; n < 3; n++) nv[n] = vertex_normal<R>(m, om, c, (ed >> (5 * n)) & 31u);
      const V3<R> n = vunit(vadd(vscale(w0, nv[0]), vadd(vscale(w1, nv[1]), vscale(w2, nv[2]))));
      const V3<R> light = {v.light[0], v.light[1], v.light[2]};
      const R fct = fmax(R(0.15), fabs(vdot(n, light)));
      r = uint8_t(fct * R(om.rgb[0]));
      gg = uint8_t(fct * R(om.rgb[1]));
      b = uint8_t(fct * R(om.rgb[2]));
    } else if (id > ntri) {
      const uint64_t f = id - ntri - 1;
      r = frag_rgb[3 * f + 0];
      gg = frag_rgb[3 * f + 1];
      b = frag_rgb[3 * f + 2];
    }
    const uint64_t o = (uint64_t(v.H - 1 - j) * uint64_t(v.W) + uint64_t(i)) * 3;
    J.rgb[o + 0] = r;
    J.rgb[o + 1] = gg;
    J.rgb[o + 2] = b;
    if (J.depth) J.depth[lp] = dep;
  }
}

// Vertex normal of the vertex on cell-local edge e of cell c, recomputed
// from the field: the cells sharing the edge in increasing linear order
// (normal.c:78-158), in each the triangles of threshold thr with a vertex
// on the edge in slot order (c_emask lists them), their smoothing weights
// summed from +0 and normalised -- the sums of vertex_normal over a stored
// mesh, bit for bit.  Lc, vc and code_c are cell c's location, corner
// values and corner code.
template <typename R>
__device__ V3<R> vertex_normal_field(const FieldDev& f, double thr, uint32_t c, const CellLoc& Lc,
                                     const double vc[8], uint32_t code_c, uint32_t e) {
  const int64_t sy = f.ncx, sz = int64_t(f.ncx) * f.ncy;
  V3<R> sum = {R(0), R(0), R(0)};
  const int nn = c_nbr.n[e];
  for (int q = 0; q < nn; q++) {
    const int dk = c_nbr.v[e][q][0], dj = c_nbr.v[e][q][1], di = c_nbr.v[e][q][2];
    const uint32_t want = uint32_t(c_nbr.v[e][q][3]);
    const bool self = (dk | dj | di) == 0;
    const uint32_t c2 = uint32_t(int64_t(c) + dk * sz + dj * sy + di);
    const CellLoc L2 = self ? Lc : locate(f, c2);
    double v2[8];
    if (self) {
#pragma unroll
      for (int n = 0; n < 8; n++) v2[n] = vc[n];
    } else {
      load_corners(f, L2, v2);
    }
    const uint32_t code = self ? code_c : corner_code(thr, v2);
    uint32_t m = c_emask[code * 19u + want];
    double g6[6] = {0, 0, 0, 0, 0, 0};
    if (m && !f.cx) cell_grid(f, L2, g6);
    for (; m; m &= m - 1u) {
      const uint64_t ent = c_cslot[code * 12u + uint32_t(__builtin_ctz(m))];
      V3<R> P[3];
      slot_positions_reg<R>(f, L2, thr, ent, v2, g6, P);
      sum = vadd(sum, smoothing_weight<R>(P));
    }
  }
  return vunit(sum);
}

// k_resolve for a frame without a stored mesh: the nearest triangle is
// found from its id (FP64: the object, cell and slot fields of tri_id64;
// FP32, a serial id: the object among the image's objects by base, the cell
// from the stored cell index, the slot from the cell store's neighbouring
// entries) and recomputed from its field.  The kernel is latency-bound (a
// chain of dependent gathers per pixel), so occupancy pays: the FP32 form
// runs at 4 waves per SIMD (128 VGPRs, a small spill), the FP64 form at 3.
template <typename R>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? 3 : 4)))
void k_resolve_field(const TessJob* __restrict__ jobs, const uint32_t* __restrict__ cell,
                     const RasterJob<R>* __restrict__ rjobs, int njobs, uint64_t npx_total,
                     const ObjMeta* __restrict__ meta, int nobj, const uint8_t* __restrict__ frag_rgb,
                     const FrameHdr* __restrict__ hdr, const uint64_t* __restrict__ key,
                     const void* __restrict__ idv) {
  (void)nobj;  // objects are found from the image's range (and, in FP64, the id)
  // a frame whose triangles outgrew the cell store is rendered again: skip it
  if (hdr->overflow & 1u) return;
  const uint64_t ntri = hdr->ntri;
  for (uint64_t p0 = uint64_t(blockIdx.x) * kThreads; p0 < npx_total; p0 += uint64_t(gridDim.x) * kThreads) {
    int jb = 0;
    if (njobs > 1) {
      int lo = 0, hi = njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rjobs[mid].px0 <= p0) lo = mid;
        else hi = mid - 1;
      }
      jb = lo;
    }
    const RasterJob<R>& J = rjobs[jb];
    const ViewDev<R> v = J.v;
    const uint64_t p = p0 + threadIdx.x;
    const uint64_t lp = p - J.px0;
    if (lp >= uint64_t(v.W) * uint64_t(v.H)) continue;
    const int32_t j = int32_t(lp / uint64_t(v.W));
    const int32_t i = int32_t(lp - uint64_t(j) * uint64_t(v.W));
    const uint64_t k = key[p];
    uint8_t r = J.bg[0], gg = J.bg[1], b = J.bg[2];
    double dep;
    // the winner: a triangle (its object, cell and slot) or a line fragment
    bool tri = false, line = false;
    uint32_t obj = 0, c = 0, sl = 0;
    uint64_t fi = 0;
    if constexpr (sizeof(R) == 8) {
      // tri_id64: (object + 1) << 40 | cell << 4 | slot
      const uint64_t id = static_cast<const uint64_t*>(idv)[p];
      dep = unord64(k);
      const uint64_t ob = id >> 40;
      line = ob == kLineObj;
      fi = id & ((uint64_t(1) << 40) - 1u);
      tri = id != 0 && !line && ob - 1u >= J.obj0 && ob - 1u < J.obj1;
      obj = uint32_t(ob - 1u);
      c = uint32_t(id >> 4);
      sl = uint32_t(id & 15u);
    } else {
      const uint32_t id = uint32_t(k);
      dep = id ? double(unord32(uint32_t(k >> 32))) : -DBL_MAX;
      line = id > ntri;
      fi = id - ntri - 1;
      if (id != 0 && id <= ntri && J.obj0 < J.obj1) {
        const uint64_t t = id - 1;
        // the winner is one of the image's own objects
        int lo = int(J.obj0), hi = int(J.obj1) - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (meta[mid].base <= t) lo = mid;
          else hi = mid - 1;
        }
        obj = uint32_t(lo);
        const ObjMeta& om = meta[lo];
        c = cell[t];
        // Within an object the cells ascend with the serial position and
        // a cell has at most 12 triangles, so the slot of t is the number
        // of the 11 positions before it, inside the object, that hold the
        // same cell.  The 11 loads are independent and share one or two
        // lines.
#pragma unroll
        for (uint32_t d = 1; d < 12; d++) sl += t >= om.base + d && cell[t - d] == c ? 1u : 0u;
        // a position outside the object can only come from a damaged cell
        // store: such a pixel keeps the background
        tri = t >= om.base && t - om.base < om.count;
      }
    }
    if (tri) {
      const ObjMeta om = meta[obj];
      const FieldDev f = jobs[om.job].f;
      double thr = jobs[om.job].ts.thr[0];
      for (uint32_t z = 1; z < kMaxThr; z++) thr = om.q == z ? jobs[om.job].ts.thr[z] : thr;
      // a cell index outside the object's field or its interior, or a slot
      // outside its cell, can only come from a damaged id or cell store:
      // such a pixel keeps the background rather than read out of bounds
      do {
        if (c >= f.ncells) break;
        const CellLoc L = locate(f, c);
        if (!interior_cell(f, L)) break;
        double cv[8];
        load_corners(f, L, cv);
        const uint32_t code = corner_code(thr, cv);
        if (sl >= c_ntri[code]) break;
        const uint64_t ent = c_cslot[code * 12u + sl];
        V3<R> wpos[3];
        slot_positions<R>(f, L, thr, ent, cv, wpos);
        Projected<R> P;
#pragma unroll
        for (int n = 0; n < 3; n++) project(v, wpos[n], P.q[n]);
        raster_box(v, P);
        R w0, w1, w2;
        pixel_weights(P, pixel_centre<R>(i, v.fw), pixel_centre<R>(j, v.fh), w0, w1, w2);
        V3<R> nv[3];
#pragma unroll 1
        for (int n = 0; n < 3; n++)
          nv[n] = vertex_normal_field<R>(f, thr, c, L, cv, code, uint32_t(ent >> (18 + 5 * n)) & 31u);
        const V3<R> n = vunit(vadd(vscale(w0, nv[0]), vadd(vscale(w1, nv[1]), vscale(w2, nv[2]))));
        const V3<R> light = {v.light[0], v.light[1], v.light[2]};
        const R fct = fmax(R(0.15), fabs(vdot(n, light)));
        r = uint8_t(fct * R(om.rgb[0]));
        gg = uint8_t(fct * R(om.rgb[1]));
        b = uint8_t(fct * R(om.rgb[2]));
      } while (false);
    } else if (line) {
      r = frag_rgb[3 * fi + 0];
      gg = frag_rgb[3 * fi + 1];
      b = frag_rgb[3 * fi + 2];
    }
    const uint64_t o = (uint64_t(v.H - 1 - j) * uint64_t(v.W) + uint64_t(i)) * 3;
    J.rgb[o + 0] = r;
    J.rgb[o + 1] = gg;
    J.rgb[o + 2] = b;
    if (J.depth) J.depth[lp] = dep;
  }
}

// ----------------------------------------------------------------------
// host launchers

int persistent_blocks() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus <= 0)
      cus = 256;
    n = ((cus * 8 + 7) / 8) * 8;
  }
  return n;
}

static uint32_t grid8(uint64_t want, uint32_t cap) {
  uint64_t g = want < cap ? want : cap;
  if (g == 0) g = 1;
  return uint32_t((g + 7) / 8 * 8);
}

hipError_t launch_count(const TessJob* jobs, int njobs, uint32_t nitems, uint32_t* cnt, uint32_t grid,
                        hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count, dim3(grid8(nitems, grid)), dim3(kThreads), 0, st, jobs, njobs, nitems, cnt);
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* cnt, uint64_t* start, uint64_t n, uint64_t* tile_tmp, hipStream_t st) {
  const uint64_t ntile = (n + kScanTile - 1) / kScanTile;
  if (ntile == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_scan_reduce, dim3(uint32_t(ntile)), dim3(kThreads), 0, st, cnt, n, tile_tmp);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, st, tile_tmp, ntile);
  hipLaunchKernelGGL(k_scan_apply, dim3(uint32_t(ntile)), dim3(kThreads), 0, st, cnt, n,
                     (const uint64_t*)tile_tmp, start);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_plan(const uint64_t* start, uint64_t ntiles, ObjMeta* meta, int nobj, RasterJob<R>* rjobs,
                       int njobs, FrameHdr* hdr, uint64_t cap, hipStream_t st) {
  hipLaunchKernelGGL(k_plan<R>, dim3(1), dim3(kThreads), 0, st, start, ntiles, meta, nobj, rjobs, njobs, hdr, cap);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_emit(const TessJob* jobs, int njobs, uint32_t nitems, const uint64_t* start, const MeshView<R>& m,
                       FrameHdr* hdr, uint32_t grid, hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit<R>, dim3(grid8(nitems, grid)), dim3(kThreads), 0, st, jobs, njobs, nitems, start, m, hdr);
  return hipGetLastError();
}

hipError_t launch_cell_starts(const uint8_t* code, const ObjMeta* meta_dev, uint32_t ntiles, uint16_t* cstart,
                              hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_starts, dim3(ntiles), dim3(kThreads), 0, st, code, meta_dev, cstart);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_vertex_normals(const MeshView<R>& m, const ObjMeta* meta_dev, const uint16_t* cstart,
                                 uint64_t count, R* out, hipStream_t st) {
  if (count == 0) return hipSuccess;
  // one item per thread (the kernel strides past 2^24 workgroups)
  const uint64_t blocks = (3 * count + kThreads - 1) / kThreads;
  const uint32_t grid = uint32_t(blocks < (uint64_t(1) << 24) ? blocks : (uint64_t(1) << 24));
  hipLaunchKernelGGL(k_smooth<R>, dim3(grid), dim3(kThreads), 0, st, m, meta_dev, cstart, out);
  return hipGetLastError();
}

hipError_t launch_zinit(uint64_t* key, void* id, bool wide, uint64_t npx, int precision, bool fresh, FrameHdr* hdr,
                        hipStream_t st) {
  const uint64_t k0 = precision == 0 ? 0x0010000000000000ull /* ord64(-DBL_MAX) */ : 0ull;
  const dim3 g(grid8((npx + 255) / 256, 4096));
  if (wide)
    hipLaunchKernelGGL(k_zinit<uint64_t>, g, dim3(256), 0, st, key, static_cast<uint64_t*>(id), npx, k0, fresh, hdr);
  else
    hipLaunchKernelGGL(k_zinit<uint32_t>, g, dim3(256), 0, st, key, static_cast<uint32_t*>(id), npx, k0, fresh, hdr);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_raster(const MeshView<R>& m, const RasterJob<R>* rjobs, int njobs, FrameHdr* hdr, int mode,
                         uint64_t* key, uint32_t* id, Frag* frags, uint32_t cap, uint32_t grid, hipStream_t st) {
  if (sizeof(R) == 4 && mode != kRasterFp32) return hipErrorInvalidValue;
  if (sizeof(R) == 8 && mode == kRasterFp32) return hipErrorInvalidValue;
  if (mode == kRasterFp32)
    hipLaunchKernelGGL((k_raster<R, kRasterFp32>), dim3(grid), dim3(kThreads), 0, st, m, rjobs, njobs, hdr, key, id,
                       frags, cap);
  else if (mode == kRasterDepth64)
    hipLaunchKernelGGL((k_raster<R, kRasterDepth64>), dim3(grid), dim3(kThreads), 0, st, m, rjobs, njobs, hdr, key,
                       id, frags, cap);
  else
    hipLaunchKernelGGL((k_raster<R, kRasterId64>), dim3(grid), dim3(kThreads), 0, st, m, rjobs, njobs, hdr, key, id,
                       frags, cap);
  return hipGetLastError();
}

hipError_t launch_frag_keys(const void* frags, bool wide, uint32_t cap, const FrameHdr* hdr, uint64_t* key,
                            uint32_t grid, hipStream_t st) {
  if (wide)
    hipLaunchKernelGGL(k_frag_keys<Frag64>, dim3(grid), dim3(256), 0, st, static_cast<const Frag64*>(frags), cap, hdr,
                       key);
  else
    hipLaunchKernelGGL(k_frag_keys<Frag>, dim3(grid), dim3(256), 0, st, static_cast<const Frag*>(frags), cap, hdr,
                       key);
  return hipGetLastError();
}

hipError_t launch_frag_ids(const void* frags, bool wide, uint32_t cap, const FrameHdr* hdr, const uint64_t* key,
                           void* id, uint32_t grid, hipStream_t st) {
  if (wide)
    hipLaunchKernelGGL(k_frag_ids<uint64_t>, dim3(grid), dim3(256), 0, st, static_cast<const Frag64*>(frags), cap,
                       hdr, key, static_cast<uint64_t*>(id));
  else
    hipLaunchKernelGGL(k_frag_ids<uint32_t>, dim3(grid), dim3(256), 0, st, static_cast<const Frag*>(frags), cap, hdr,
                       key, static_cast<uint32_t*>(id));
  return hipGetLastError();
}

hipError_t launch_line_frags(const uint32_t* pix, const double* depth, uint32_t cap, const FrameHdr* hdr, int pass,
                             int precision, uint64_t* key, void* id, bool wide, hipStream_t st) {
  if (cap == 0) return hipSuccess;
  const dim3 g((cap + 255) / 256);
  if (wide)
    hipLaunchKernelGGL(k_line_frags<uint64_t>, g, dim3(256), 0, st, pix, depth, hdr, pass, precision, key,
                       static_cast<uint64_t*>(id));
  else
    hipLaunchKernelGGL(k_line_frags<uint32_t>, g, dim3(256), 0, st, pix, depth, hdr, pass, precision, key,
                       static_cast<uint32_t*>(id));
  return hipGetLastError();
}

template <typename R>
hipError_t launch_resolve(const MeshView<R>& m, const RasterJob<R>* rjobs, int njobs, uint64_t npx_total,
                          const ObjMeta* meta, int nobj, const uint8_t* frag_rgb, const FrameHdr* hdr,
                          const uint64_t* key, const uint32_t* id, uint32_t grid, hipStream_t st) {
  hipLaunchKernelGGL(k_resolve<R>, dim3(grid8((npx_total + kThreads - 1) / kThreads, grid)), dim3(kThreads), 0, st,
                     m, rjobs, njobs, npx_total, meta, nobj, frag_rgb, hdr, key, id);
  return hipGetLastError();
}

hipError_t launch_mesh_compare(const MeshView<double>& a, const ObjMeta* ma, const void* b, int b_fp32,
                               const ObjMeta* mb, uint64_t count, unsigned long long* out, hipStream_t st) {
  if (count == 0) return hipSuccess;
  const uint32_t grid = grid8((count + kThreads - 1) / kThreads, uint32_t(persistent_blocks()) * 4);
  if (b_fp32)
    hipLaunchKernelGGL((k_mesh_compare<double, float>), dim3(grid), dim3(kThreads), 0, st, a, ma,
                       *static_cast<const MeshView<float>*>(b), mb, count, out);
  else
    hipLaunchKernelGGL((k_mesh_compare<double, double>), dim3(grid), dim3(kThreads), 0, st, a, ma,
                       *static_cast<const MeshView<double>*>(b), mb, count, out);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_tess_raster(const TessJob* jobs, int njobs, uint32_t nitems, const uint64_t* start, uint32_t* cell,
                              uint64_t cap, const RasterJob<R>* rjobs, const uint32_t* jimg, FrameHdr* hdr, int mode,
                              uint64_t* key, void* id, void* frags, uint32_t fcap, uint32_t grid, hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  const dim3 g(grid8(nitems, grid));
  if (mode == kRasterAppend) {
    hipLaunchKernelGGL((k_tess_raster<R, kRasterAppend>), g, dim3(kThreads), 0, st, jobs, njobs, nitems, start, cell,
                       cap, rjobs, jimg, hdr, key, id, frags, fcap);
  } else if constexpr (sizeof(R) == 4) {
    if (mode != kRasterFp32) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_tess_raster<R, kRasterFp32>), g, dim3(kThreads), 0, st, jobs, njobs, nitems, start, cell, cap,
                       rjobs, jimg, hdr, key, id, frags, fcap);
  } else {
    if (mode == kRasterKey64)
      hipLaunchKernelGGL((k_tess_raster<R, kRasterKey64>), g, dim3(kThreads), 0, st, jobs, njobs, nitems, start, cell,
                         cap, rjobs, jimg, hdr, key, id, frags, fcap);
    else if (mode == kRasterId64)
      hipLaunchKernelGGL((k_tess_raster<R, kRasterId64>), g, dim3(kThreads), 0, st, jobs, njobs, nitems, start, cell,
                         cap, rjobs, jimg, hdr, key, id, frags, fcap);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_resolve_field(const TessJob* jobs, const uint32_t* cell, const RasterJob<R>* rjobs, int njobs,
                                uint64_t npx_total, const ObjMeta* meta, int nobj, const uint8_t* frag_rgb,
                                const FrameHdr* hdr, const uint64_t* key, const void* id, hipStream_t st) {
  hipLaunchKernelGGL(k_resolve_field<R>, dim3(grid8((npx_total + kThreads - 1) / kThreads, 1u << 30)), dim3(kThreads),
                     0, st, jobs, cell, rjobs, njobs, npx_total, meta, nobj, frag_rgb, hdr, key, id);
  return hipGetLastError();
}

template hipError_t launch_tess_raster<double>(const TessJob*, int, uint32_t, const uint64_t*, uint32_t*, uint64_t,
                                               const RasterJob<double>*, const uint32_t*, FrameHdr*, int, uint64_t*,
                                               void*, void*, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_tess_raster<float>(const TessJob*, int, uint32_t, const uint64_t*, uint32_t*, uint64_t,
                                              const RasterJob<float>*, const uint32_t*, FrameHdr*, int, uint64_t*,
                                              void*, void*, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_resolve_field<double>(const TessJob*, const uint32_t*, const RasterJob<double>*, int,
                                                 uint64_t, const ObjMeta*, int, const uint8_t*, const FrameHdr*,
                                                 const uint64_t*, const void*, hipStream_t);
template hipError_t launch_resolve_field<float>(const TessJob*, const uint32_t*, const RasterJob<float>*, int,
                                                uint64_t, const ObjMeta*, int, const uint8_t*, const FrameHdr*,
                                                const uint64_t*, const void*, hipStream_t);
template hipError_t launch_plan<double>(const uint64_t*, uint64_t, ObjMeta*, int, RasterJob<double>*, int, FrameHdr*,
                                        uint64_t, hipStream_t);
template hipError_t launch_plan<float>(const uint64_t*, uint64_t, ObjMeta*, int, RasterJob<float>*, int, FrameHdr*,
                                       uint64_t, hipStream_t);
template hipError_t launch_emit<double>(const TessJob*, int, uint32_t, const uint64_t*, const MeshView<double>&,
                                        FrameHdr*, uint32_t, hipStream_t);
template hipError_t launch_emit<float>(const TessJob*, int, uint32_t, const uint64_t*, const MeshView<float>&,
                                       FrameHdr*, uint32_t, hipStream_t);
template hipError_t launch_vertex_normals<double>(const MeshView<double>&, const ObjMeta*, const uint16_t*,
                                                  uint64_t, double*, hipStream_t);
template hipError_t launch_vertex_normals<float>(const MeshView<float>&, const ObjMeta*, const uint16_t*, uint64_t,
                                                 float*, hipStream_t);
template hipError_t launch_raster<double>(const MeshView<double>&, const RasterJob<double>*, int, FrameHdr*, int,
                                          uint64_t*, uint32_t*, Frag*, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_raster<float>(const MeshView<float>&, const RasterJob<float>*, int, FrameHdr*, int,
                                         uint64_t*, uint32_t*, Frag*, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_resolve<double>(const MeshView<double>&, const RasterJob<double>*, int, uint64_t,
                                           const ObjMeta*, int, const uint8_t*, const FrameHdr*, const uint64_t*,
                                           const uint32_t*, uint32_t, hipStream_t);
template hipError_t launch_resolve<float>(const MeshView<float>&, const RasterJob<float>*, int, uint64_t,
                                          const ObjMeta*, int, const uint8_t*, const FrameHdr*, const uint64_t*,
                                          const uint32_t*, uint32_t, hipStream_t);

hipError_t upload_tables(const NbrTable& nt) {
  // case table of triangulate.c:204-270: reduced mask -> triangles as
  // (tail, head) tetra-vertex pairs
  static const int ntri[8] = {0, 1, 1, 2, 1, 2, 2, 1};
  static const int pairs[8][2][3][2] = {
      {{{0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}}},
      {{{0, 1}, {0, 3}, {0, 2}}, {{0, 0}, {0, 0}, {0, 0}}},
      {{{1, 0}, {1, 2}, {1, 3}}, {{0, 0}, {0, 0}, {0, 0}}},
      {{{0, 2}, {1, 3}, {0, 3}}, {{0, 2}, {1, 2}, {1, 3}}},
      {{{2, 0}, {2, 3}, {2, 1}}, {{0, 0}, {0, 0}, {0, 0}}},
      {{{2, 1}, {0, 3}, {2, 3}}, {{2, 1}, {0, 1}, {0, 3}}},
      {{{0, 2}, {2, 3}, {1, 3}}, {{0, 2}, {1, 3}, {0, 1}}},
      {{{3, 0}, {3, 2}, {3, 1}}, {{0, 0}, {0, 0}, {0, 0}}},
  };
  static const int tet[6][4] = {{0, 7, 3, 1}, {0, 7, 1, 5}, {0, 7, 5, 4},
                                {0, 7, 4, 6}, {0, 7, 6, 2}, {0, 7, 2, 3}};
  hipError_t e;
  // per corner code: triangle count and every slot's corners and edge ids
  static uint8_t nt8[256];
  static uint64_t cs[256 * 12];
  static uint16_t em[256 * 19];
  auto eid = [](uint32_t a, uint32_t b) -> uint32_t {
    const uint32_t lo = a < b ? a : b;
    switch (a ^ b) {
      case 1: return lo >> 1;
      case 2: return 4 + (lo & 1) + 2 * (lo >> 2);
      case 4: return 8 + (lo & 3);
      case 6: return 12 + (lo & 1);
      case 5: return 14 + ((lo >> 1) & 1);
      case 3: return 16 + (lo >> 2);
      default: return 18;
    }
  };
  for (int code = 0; code < 256; code++) {
    int sl = 0;
    for (int t = 0; t < 6; t++) {
      int mk = 0;
      for (int n = 0; n < 4; n++) mk |= ((code >> tet[t][n]) & 1) << n;
      const bool rev = mk >= 8;
      const int mm = rev ? 15 - mk : mk;
      for (int s = 0; s < ntri[mm]; s++, sl++) {
        uint64_t ent = 0;
        for (int v = 0; v < 3; v++) {
          const int q = rev ? 2 - v : v;
          const uint32_t tail = uint32_t(tet[t][pairs[mm][s][q][0]]);
          const uint32_t head = uint32_t(tet[t][pairs[mm][s][q][1]]);
          ent |= uint64_t(tail | head << 3) << (6 * v);
          ent |= uint64_t(eid(tail, head)) << (18 + 5 * v);
          em[code * 19 + eid(tail, head)] |= uint16_t(1u << sl);
        }
        cs[code * 12 + sl] = ent;
      }
    }
    nt8[code] = uint8_t(sl);
  }
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_ntri), nt8, sizeof(nt8));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_cslot), cs, sizeof(cs));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_emask), em, sizeof(em));
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_nbr), &nt, sizeof(nt));
}

NbrTable build_nbr_table() {
  // every cell within one step holding both endpoints of edge e as a
  // tetra edge, in increasing (dk, dj, di) = increasing linear order
  static const int tet[6][4] = {{0, 7, 3, 1}, {0, 7, 1, 5}, {0, 7, 5, 4},
                                {0, 7, 4, 6}, {0, 7, 6, 2}, {0, 7, 2, 3}};
  auto is_tet_edge = [&](int a, int b) {
    if (a == b) return false;
    for (int t = 0; t < 6; t++) {
      bool ha = false, hb = false;
      for (int n = 0; n < 4; n++) {
        ha |= tet[t][n] == a;
        hb |= tet[t][n] == b;
      }
      if (ha && hb) return true;
    }
    return false;
  };
  auto eid = [](int a, int b) {
    const int lo = a < b ? a : b;
    switch (a ^ b) {
      case 1: return lo >> 1;
      case 2: return 4 + (lo & 1) + 2 * (lo >> 2);
      case 4: return 8 + (lo & 3);
      case 6: return 12 + (lo & 1);
      case 5: return 14 + ((lo >> 1) & 1);
      case 3: return 16 + (lo >> 2);
      default: return 18;
    }
  };
  NbrTable nt = {};
  for (int a = 0; a < 8; a++) {
    for (int b = a + 1; b < 8; b++) {
      if (!is_tet_edge(a, b)) continue;
      const int e = eid(a, b);
      const int A[3] = {a & 1, (a >> 1) & 1, a >> 2};
      const int B[3] = {b & 1, (b >> 1) & 1, b >> 2};
      for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            const int d[3] = {di, dj, dk};
            int la[3], lb[3];
            bool ok = true;
            for (int x = 0; x < 3; x++) {
              la[x] = A[x] - d[x];
              lb[x] = B[x] - d[x];
              ok = ok && la[x] >= 0 && la[x] <= 1 && lb[x] >= 0 && lb[x] <= 1;
            }
            if (!ok) continue;
            const int ca = la[0] + 2 * la[1] + 4 * la[2];
            const int cb = lb[0] + 2 * lb[1] + 4 * lb[2];
            if (!is_tet_edge(ca, cb)) continue;
            const int q = nt.n[e]++;
            nt.v[e][q][0] = int8_t(dk);
            nt.v[e][q][1] = int8_t(dj);
            nt.v[e][q][2] = int8_t(di);
            nt.v[e][q][3] = int8_t(eid(ca, cb));
          }
    }
  }
  return nt;
}

}  // namespace c3d
