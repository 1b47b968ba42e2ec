// Internal declarations shared by the device kernels (c3d_kernels.hip) and
// the host orchestration behind the C ABI (c3d_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "c3d_geom.h"
#include "c3d_tiles.h"

namespace c3d {

constexpr int kThreads = 256;           // one workgroup = 4 wave64
constexpr int kTileChunks = kTileCells / kThreads;  // chunks of 256 cells per tile
constexpr int kMaxThr = 4;              // thresholds classified per field read
constexpr int kScanTile = 1024;         // entries per workgroup of the tile-count scan

// One FP64 field on the device plus its lattice coordinates: either the
// identity map of the three grids, or an SoA coordinate table (cx != null).
struct FieldDev {
  const double* f;
  const double* gx;
  const double* gy;
  const double* gz;
  const double* cx;
  const double* cy;
  const double* cz;
  uint32_t nx, ny, nz;            // lattice points
  uint32_t ncx, ncy, ncz;         // cells
  uint32_t ncells;
  uint32_t ntiles;                // ncz * tg.cols
  uint32_t nitems;                // tg.cols * tg.nzc
  FastDiv div_ncx, div_ncy;
  TileGeo tg;
};

// Thresholds classified in one pass over a field, each with the offset of
// its object's first tile in the frame's tile-count array.
struct ThrSet {
  double thr[kMaxThr];
  uint32_t boff[kMaxThr];
  uint32_t obj[kMaxThr];  // the frame's contour object of each threshold
  int n;
};

// One pass over one field: the unit of work of the tessellation kernels.
// Jobs are stored in a device table; their work items are numbered
// consecutively from item0.  xoff is the sum of ceil(items / 8) over the
// earlier jobs: the job's first position in each XCD's walk over the
// launch's work items.
struct TessJob {
  FieldDev f;
  ThrSet ts;
  uint32_t item0;
  uint32_t xoff;
  uint64_t code_off;  // a stored mesh: threshold q's cell codes at code_off + q * f.ncells
};

// One stored triangle, as two records in two arrays, both in serial order:
// its positions, and what the smoothing gathers from neighbouring
// triangles (the weight and the cell), so that those gathers read 32-byte
// (FP64) or 16-byte (FP32) records.  The edge word sits where it keeps both
// records a multiple of 8 bytes (a drain stores them as 8-byte words):
//   FP64: MeshPos 72 B {p[9]},        MeshAux 32 B {w[3], cell, edges}
//   FP32: MeshPos 40 B {p[9], edges}, MeshAux 16 B {w[3], cell}
// p: vertex v, axis d at 3v + d; w: area * unit face normal, the smoothing
// weight; cell: linear cell index within its field; edges: three 5-bit
// cell-local lattice edge ids | interior << 15.
template <typename R>
struct MeshPos;
template <>
struct alignas(8) MeshPos<double> {
  double p[9];
};
template <>
struct alignas(8) MeshPos<float> {
  float p[9];
  uint32_t edges;
};
template <typename R>
struct MeshAux;
template <>
struct alignas(8) MeshAux<double> {
  double w[3];
  uint32_t cell;
  uint32_t edges;
};
template <>
struct alignas(8) MeshAux<float> {
  float w[3];
  uint32_t cell;
};
static_assert(sizeof(MeshPos<double>) == 72 && sizeof(MeshAux<double>) == 32 && sizeof(MeshPos<float>) == 40 &&
                  sizeof(MeshAux<float>) == 16,
              "record layout");

// Triangle storage: cap records of each kind, and the corner code (8 bits)
// of every cell of every object's field, object after object
// (ObjMeta::code_off).
template <typename R>
struct MeshView {
  MeshPos<R>* pos;
  MeshAux<R>* aux;
  uint64_t cap;
  uint8_t* code;
};

// Per contour object, on the device.
struct ObjMeta {
  uint64_t base;           // first triangle of the object in the mesh
  uint64_t count;
  const uint64_t* bstart;  // triangle start of each tile of the object's field (ntiles + 1)
  uint32_t ncx, ncy, ncz;
  FastDiv div_ncx, div_ncy;
  TileGeo tg;              // the tiles of the object's field
  uint32_t tile_off;       // offset of the object's first tile in the frame's tile arrays
  uint32_t job, q;         // its field pass and threshold slot there
  uint8_t rgb[4];
  uint64_t code_off;       // a stored mesh: first of its cells' corner codes (and cell starts)
};

// One covered (pixel, triangle) of the stored-mesh FP64 z-buffer pass.
struct Frag {
  uint32_t px;
  uint32_t id;
  uint64_t key;
};

// A triangle of an FP64 frame is named by tri_id64: (object + 1) << 40 |
// cell << 4 | slot.  These ids ascend in the reference's drawing order
// (contour, then cell index k, j, i, then slot; contour/main.c:49-112), so
// "largest id" breaks depth ties as its sequential >= test does, and they
// need no count pass: the resolve reads the object, cell and slot straight
// from the id.  Line fragments follow every object (kLineObj).
constexpr uint64_t kLineObj = 0xffffffull;
__host__ __device__ inline uint64_t tri_id64(uint32_t obj, uint32_t cell, uint32_t slot) {
  return (uint64_t(obj) + 1u) << 40 | uint64_t(cell) << 4 | slot;
}

// One covered (pixel, triangle) of a fused FP64 frame.
struct Frag64 {
  uint64_t key;
  uint64_t id;  // tri_id64
  uint32_t px;
  uint32_t pad;
};

// Cells sharing each of the 19 cell-local lattice edges, as (dk, dj, di,
// local id there) in increasing linear order (normal.c:78-158).
struct NbrTable {
  int8_t v[19][4][4];
  uint8_t n[19];
};
NbrTable build_nbr_table();

// One image of a frame: its camera, its contour objects [obj0, obj1) (their
// triangles are contiguous in the mesh; images showing the same objects
// share them), its pixels in the frame's z-buffer from px0, and where its
// colours (and depths) go.  The raster kernels walk a work space that
// concatenates every image's triangle range: image j owns work items
// [wk0, wk0 + tri1 - tri0).
template <typename R>
struct RasterJob {
  ViewDev<R> v;
  uint64_t tri0, tri1;   // filled on the device by k_plan
  uint64_t wk0;          // filled on the device by k_plan
  uint64_t px0;
  uint32_t obj0, obj1;
  uint8_t* rgb;          // device, W*H*3 in PPM row order
  double* depth;         // device or null
  uint8_t bg[4];
};

// Frame-wide values on the device.  k_plan writes ntri, nwork and the
// overflow bit 1 (meshes, FP32 frames); k_zinit clears nfrag and bit 2, and
// in an FP64 frame, which has no plan, also ntri and bit 1, and its Append
// pass then adds up ntri; the host writes nline before the plan runs.
struct FrameHdr {
  uint64_t ntri;        // triangles in the mesh (or the frame)
  uint64_t nwork;       // raster work items (sum of the images' triangle ranges)
  uint32_t overflow;    // bit 1: mesh capacity too small; bit 2: FP64 fragment list overflowed
  uint32_t nfrag;       // FP64 fragments kept by the depth pass
  uint32_t nline;       // line fragments of the frame
  uint32_t pad;
};

// ----------------------------------------------------------------------
// launchers (c3d_kernels.hip); all enqueue on `st` and return hipError_t

hipError_t upload_tables(const NbrTable& nt);
int persistent_blocks();  // workgroups that fill the device once (multiple of 8)

hipError_t launch_count(const TessJob* jobs, int njobs, uint32_t nitems, uint32_t* cnt, uint32_t grid,
                        hipStream_t st);
hipError_t launch_scan(const uint32_t* cnt, uint64_t* start, uint64_t n, uint64_t* tile_tmp,
                       hipStream_t st);
// object bases/counts, frame total and per-image triangle ranges from the
// tile starts (one workgroup)
template <typename R>
hipError_t launch_plan(const uint64_t* start, uint64_t ntiles, ObjMeta* meta, int nobj, RasterJob<R>* rjobs,
                       int njobs, FrameHdr* hdr, uint64_t cap, hipStream_t st);
template <typename R>
hipError_t launch_emit(const TessJob* jobs, int njobs, uint32_t nitems, const uint64_t* start,
                       const MeshView<R>& m, FrameHdr* hdr, uint32_t grid, hipStream_t st);
// every cell's first triangle within its tile (cstart, 16 bits), for one
// object of a stored mesh, from the cell codes the emit pass stored
hipError_t launch_cell_starts(const uint8_t* code, const ObjMeta* meta_dev, uint32_t ntiles, uint16_t* cstart,
                              hipStream_t st);
// unit vertex normals of every triangle of one object, 9 values of the
// mesh's precision per triangle (NaN rows for triangles outside the
// rendered interior); needs the object's cell starts
template <typename R>
hipError_t launch_vertex_normals(const MeshView<R>& m, const ObjMeta* meta_dev, const uint16_t* cstart,
                                 uint64_t count, R* out, hipStream_t st);
// FP64 mesh a against mesh b (FP64, or FP32 when b_fp32; b points to its
// MeshView) for one object: out[0..4] as k_mesh_compare documents
hipError_t launch_mesh_compare(const MeshView<double>& a, const ObjMeta* ma, const void* b, int b_fp32,
                               const ObjMeta* mb, uint64_t count, unsigned long long* out, hipStream_t st);
// Frames without a stored mesh: tessellation fused with the raster.  Every
// triangle is drawn from registers into each image that shows its object
// (jimg: per-field-pass offsets [njobs + 1] into a list of the images that
// show any of the pass's objects).  FP64 frames name triangles by
// tri_id64 (id: uint64_t per pixel, frags: Frag64) and need no tile starts
// (start, cell: null); their Append pass adds its triangles to hdr->ntri.
// FP32 frames number them serially from the tile starts (id: uint32_t,
// frags: Frag) and store each one's cell index (cell[cap]).  `mode` is a
// raster mode.
template <typename R>
hipError_t launch_tess_raster(const TessJob* jobs, int njobs, uint32_t nitems, const uint64_t* start, uint32_t* cell,
                              uint64_t cap, const RasterJob<R>* rjobs, const uint32_t* jimg, FrameHdr* hdr, int mode,
                              uint64_t* key, void* id, void* frags, uint32_t fcap, uint32_t grid, hipStream_t st);
// resolve of a fused frame: the nearest triangles and their neighbours'
// smoothing weights are recomputed from the fields (id as launch_tess_raster)
template <typename R>
hipError_t launch_resolve_field(const TessJob* jobs, const uint32_t* cell, const RasterJob<R>* rjobs, int njobs,
                                uint64_t npx_total, const ObjMeta* meta, int nobj, const uint8_t* frag_rgb,
                                const FrameHdr* hdr, const uint64_t* key, const void* id, hipStream_t st);
// clears the z-buffer (and the ids: uint64_t per pixel when wide, else
// uint32_t), the header's fragment count and list-overflow bit, and with
// fresh (FP64 frames) also its triangle count and capacity bit
hipError_t launch_zinit(uint64_t* key, void* id, bool wide, uint64_t npx, int precision, bool fresh, FrameHdr* hdr,
                        hipStream_t st);
// raster modes:
//   kRasterFp32     FP32 z-test over packed (depth << 32 | id) keys
//   kRasterDepth64  FP64 z-test keeping the fragments that may end up nearest
//   kRasterId64     FP64 id pass over the geometry (runs only when the
//                   fragment list overflowed)
//   kRasterAppend   no z-test: every covered pixel goes to the fragment list
//                   with its key (FP64: the depth's order key; FP32: the
//                   packed key); launch_frag_keys then applies the list
//   kRasterKey64    FP64 z-test without a list (runs only when the list
//                   overflowed, before the geometry id pass)
// The fused kernel takes kRasterAppend and, after an overflow, kRasterFp32
// or kRasterKey64 + kRasterId64; the stored-mesh raster takes the first three.
enum { kRasterFp32 = 0, kRasterDepth64 = 1, kRasterId64 = 2, kRasterAppend = 3, kRasterKey64 = 4 };
template <typename R>
hipError_t launch_raster(const MeshView<R>& m, const RasterJob<R>* rjobs, int njobs, FrameHdr* hdr, int mode,
                         uint64_t* key, uint32_t* id, Frag* frags, uint32_t cap, uint32_t grid, hipStream_t st);
// z-buffer keys of every listed fragment (reads the count from hdr; does
// nothing when the list overflowed); wide: Frag64, else Frag
hipError_t launch_frag_keys(const void* frags, bool wide, uint32_t cap, const FrameHdr* hdr, uint64_t* key,
                            uint32_t grid, hipStream_t st);
// FP64 id pass over the fragment list (reads the count from hdr; does
// nothing when the list overflowed); wide: Frag64 and uint64_t ids
hipError_t launch_frag_ids(const void* frags, bool wide, uint32_t cap, const FrameHdr* hdr, const uint64_t* key,
                           void* id, uint32_t grid, hipStream_t st);
// line fragments [0, hdr->nline) of the frame; `cap` sizes the grid; wide:
// uint64_t ids after every object (kLineObj), else uint32_t ids after the
// frame's hdr->ntri triangles
hipError_t launch_line_frags(const uint32_t* pix, const double* depth, uint32_t cap, const FrameHdr* hdr,
                             int pass, int precision, uint64_t* key, void* id, bool wide, hipStream_t st);
template <typename R>
hipError_t launch_resolve(const MeshView<R>& m, const RasterJob<R>* rjobs, int njobs, uint64_t npx_total,
                          const ObjMeta* meta, int nobj, const uint8_t* frag_rgb, const FrameHdr* hdr,
                          const uint64_t* key, const uint32_t* id, uint32_t grid, hipStream_t st);

}  // namespace c3d
