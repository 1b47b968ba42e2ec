// Internal declarations shared by the device kernels (c3d_kernels.hip) and
// the host orchestration behind the C ABI (c3d_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace c3d {

constexpr int kThreads = 256;           // one workgroup = 4 wave64
constexpr int kCellsPerBlock = 1024;    // tessellation tile: 4 chunks of 256 cells
constexpr int kMaxThr = 4;              // thresholds classified per field read

// Division by a runtime constant with a multiply-high (Granlund-Montgomery,
// "round-up" variant): q = (umulhi(n, m) + n) >> s for every 32-bit n.
struct FastDiv {
  uint32_t d, m, s;
};
FastDiv make_fastdiv(uint32_t d);

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
  uint32_t nblocks;               // ceil(ncells / kCellsPerBlock)
  FastDiv div_ncx, div_ncy;
};

// Thresholds classified in one pass over a field, each with the offset of
// its object's first tile in the global tile-count array.
struct ThrSet {
  double thr[kMaxThr];
  uint32_t boff[kMaxThr];
  int n;
};

// Triangle storage, structure of arrays with capacity cap: component c of
// triangle t lives at [c * cap + t].  R is double (FP64 form) or float
// (FP32 form).
template <typename R>
struct MeshView {
  R* pos;          // [9][cap]: vertex v, axis d at (3v + d)
  R* wfn;          // [3][cap]: area * unit face normal, the smoothing weight
  uint32_t* cell;  // linear cell index within its field
  uint32_t* edges; // three 5-bit cell-local lattice edge ids | interior << 15
  uint64_t cap;
};

// Per contour object, on the device.
struct ObjMeta {
  uint64_t base;           // first triangle of the object in the mesh
  uint64_t count;
  const uint64_t* bstart;  // triangle start of each tile of the object's field (nblocks + 1)
  uint32_t ncx, ncy, ncz;
  uint8_t rgb[4];
};

// One covered (pixel, triangle) of the FP64 z-buffer pass.
struct Frag {
  uint32_t px;
  uint32_t id;
  uint64_t key;
};

// Cells sharing each of the 19 cell-local lattice edges, as (dk, dj, di,
// local id there) in increasing linear order (normal.c:78-158).
struct NbrTable {
  int8_t v[19][4][4];
  uint8_t n[19];
};
NbrTable build_nbr_table();

// Camera/screen constants; nc = n.camera, d = n.center, lxx = lx.lx,
// lyy = ly.ly are formed on the host in FP64 exactly as project.c:34-36,
// 62-63 form them per call.
template <typename R>
struct ViewDev {
  R cam[3], ctr[3], lx[3], ly[3], nrm[3], light[3];
  R nc, d, lxx, lyy;
  R fw, fh;          // width, height as reals
  int32_t W, H;
};

// ----------------------------------------------------------------------
// launchers (c3d_kernels.hip); all enqueue on `st` and return hipError_t

hipError_t upload_tables(const NbrTable& nt);
hipError_t launch_count(const FieldDev& f, const ThrSet& ts, uint32_t* cnt, hipStream_t st);
hipError_t launch_scan(const uint32_t* cnt, uint64_t* start, uint64_t n, uint64_t* tile_tmp,
                       hipStream_t st);
hipError_t launch_gather_u64(const uint64_t* src, const uint32_t* idx, uint64_t* dst, int n,
                             hipStream_t st);
template <typename R>
hipError_t launch_emit(const FieldDev& f, const ThrSet& ts, const uint64_t* start,
                       const MeshView<R>& m, hipStream_t st);
// unit vertex normals of every triangle of one object, 9 doubles per
// triangle (NaN rows for triangles outside the rendered interior)
template <typename R>
hipError_t launch_vertex_normals(const MeshView<R>& m, const ObjMeta* meta_dev, uint64_t count,
                                 double* out, hipStream_t st);
hipError_t launch_zinit(uint64_t* key, uint32_t* id, uint64_t npx, int precision, hipStream_t st);
template <typename R>
hipError_t launch_raster(const MeshView<R>& m, uint64_t ntri, const ViewDev<R>& v, int pass,
                         uint64_t* key, uint32_t* id, hipStream_t st);
// FP64 depth pass that keeps the fragments that may end up nearest; the
// count goes to counter[0], counter[1] is set on overflow
hipError_t launch_raster_frags(const MeshView<double>& m, uint64_t ntri, const ViewDev<double>& v,
                               uint64_t* key, Frag* frags, uint32_t cap, uint32_t* counter,
                               hipStream_t st);
// FP64 id pass over the fragment list
hipError_t launch_frag_ids(const Frag* frags, uint32_t n, const uint64_t* key, uint32_t* id,
                           hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st);
hipError_t launch_line_frags(const uint32_t* pix, const double* depth, uint32_t nfrag,
                             uint32_t id_base, int pass, int precision, uint64_t* key, uint32_t* id,
                             hipStream_t st);
template <typename R>
hipError_t launch_resolve(const MeshView<R>& m, uint64_t ntri, const ViewDev<R>& v,
                          const ObjMeta* meta, int nobj, const uint8_t* frag_rgb, const uint8_t bg[3],
                          const uint64_t* key, const uint32_t* id, uint8_t* rgb, double* depth,
                          hipStream_t st);

}  // namespace c3d
