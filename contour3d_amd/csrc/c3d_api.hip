// C ABI of the MI355X Contour3D library: contour3d_execute (contour3d.h) and
// the extensions of contour3d_amd.h.  Host orchestration only; the compute
// runs in the kernels of c3d_kernels.hip on one library stream.
#include "contour3d.h"
#include "contour3d_amd.h"
#include "c3d_internal.h"
#include "c3d_wait.h"

#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

// Device memory of contour3d_amd_execute_rgb_noalloc: one static array,
// which every GPU the library runs on holds until the process ends.  Host
// code never passes its name to a kernel: set-up (contour3d_amd_init)
// takes its address on the current device with hipGetSymbolAddress, and
// the frame's buffers are regions of that address.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace c3d {
namespace {

// "[CONTOUR3D ERROR] " + message on stderr, the reference's convention
// (src/contour3d/logger.c:5-18)
void log_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "[CONTOUR3D ERROR] ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
  va_end(ap);
}

#define C3D_HIP(expr)                                                                  \
  do {                                                                                 \
    const hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                            \
      log_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

// Failure injection for the error contract (CONTOUR3D_AMD_TUNE_FAIL_ALLOC):
// while positive, every device or pinned allocation counts it down, and the
// allocation that brings it from 1 to 0 fails as an out-of-memory would.
std::atomic<int64_t> g_fail_alloc{0};

bool injected_alloc_failure() {
  int64_t v = g_fail_alloc.load();
  while (v > 0) {
    if (g_fail_alloc.compare_exchange_weak(v, v - 1)) return v == 1;
  }
  return false;
}

// grow-only device buffer, or (fixed) a region of the static scratch array
// that never grows
// A number no device buffer allocation has had before (DevBuf::gen).
uint64_t next_buf_gen() {
  static std::atomic<uint64_t> g{0};
  return ++g;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  uint64_t gen = 0;  // the current allocation's number (0: none)
  ~DevBuf() = default;
  // grows to at least `want` bytes, with a quarter of headroom unless
  // `exact` (a growing workload reallocates rarely; each reallocation frees
  // first, which waits for the device)
  hipError_t ensure(size_t want, bool exact = false) {
    if (want <= bytes) return hipSuccess;
    if (fixed) return hipErrorOutOfMemory;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    size_t n = exact ? want : want + want / 4 + 256;
    if (injected_alloc_failure()) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      gen = 0;
      return e;
    }
    bytes = n;
    gen = next_buf_gen();
    return hipSuccess;
  }
  void release() {
    if (fixed) return;  // the scratch array lives as long as the process
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    gen = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// pinned host staging buffer; a fixed one is sized once, at set-up
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (fixed) return hipErrorOutOfMemory;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    if (injected_alloc_failure()) return hipErrorOutOfMemory;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// lattice coordinates of one (converter, grids) geometry
// geometry cache limits (entries, bytes of converter coordinate tables)
constexpr size_t kMaxGeoms = 1024;
constexpr size_t kGeomTableBytes = size_t(16) << 30;

struct Geom {
  void* conv = nullptr;
  const double* g[3] = {nullptr, nullptr, nullptr};
  size_t n[3] = {0, 0, 0};
  uint64_t hash = 0;
  std::vector<double> host;  // the grid values (nx + ny + nz), compared on a hash match
  bool identity = true;
  DevBuf grids;  // nx + ny + nz doubles
  DevBuf table;  // 3 * npts doubles when not the identity
  uint64_t stamp = 0;
  std::atomic<int> busy{0};  // frames in flight that read it (never evicted then)
};

struct Frags {
  std::vector<uint32_t> pix;
  std::vector<double> depth;
  std::vector<uint8_t> rgb;
};

struct Prof {
  bool on = false;
  double ms[CONTOUR3D_AMD_NSTAGES] = {};
  uint64_t calls[CONTOUR3D_AMD_NSTAGES] = {};
  uint64_t ntri = 0;
  struct Pend {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pend> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t cur_a = nullptr;
  int cur_stage = -1;
  hipEvent_t in_a = nullptr;  // a stage timed inside the current one
  int in_stage = -1;
};

}  // namespace
}  // namespace c3d

struct contour3d_amd_mesh {
  int precision = 0;
  int precision_built = -1;  // form the buffers were sized for
  c3d::DevBuf pos;     // a stored mesh: cap position records (c3d::MeshPos)
  c3d::DevBuf aux;     // and cap gather records (c3d::MeshAux)
  c3d::DevBuf cell;    // a frame: the cell index of each triangle (cap entries)
  c3d::DevBuf bstart;  // per-tile triangle starts of every object's field
  c3d::DevBuf meta;    // ObjMeta per object
  c3d::DevBuf code;    // a stored mesh: the corner code of every cell of every object (ObjMeta::code_off)
  // and each cell's first triangle within its tile: a cache made by the
  // first smoothing of a build (calls on a mesh hold the library's lock)
  mutable c3d::DevBuf cstart;
  mutable bool starts_ready = false;
  uint64_t code_bytes = 0;
  uint64_t cap = 0;
  uint64_t total = 0;
  uint64_t ntiles = 0;  // entries of bstart minus one
  std::vector<uint64_t> base, count;
  std::vector<uint32_t> tiles;  // tiles of each object's field
  std::vector<c3d::TileGeo> geo;  // and their layout
  std::vector<contour3d_color_t> color;
  void release() {
    pos.release();
    aux.release();
    cell.release();
    bstart.release();
    meta.release();
    code.release();
    cstart.release();
    starts_ready = false;
    cap = 0;
  }
};

namespace c3d {
namespace {

// initial fragment-list entries of an FP64 frame (24 B each; 16 B for a
// stored mesh): every covered pixel of a 512^3 field's frame at 1920x1080
// fits
constexpr uint32_t kDefaultFragCap = 1u << 24;

// Host-side layout of a frame's tessellation.
struct TessLayout {
  uint64_t nbt = 0;  // tile-count entries (the objects' tiles)
  uint32_t ntj = 0;  // work items of the field passes
  int njobs = 0;     // field passes
  int nobj = 0;
};

// Stream, buffers and host staging of one frame.  The synchronous entry
// points use the library's main frame; the serving loop alternates two, so
// that one frame's host work and device work overlap the other's.
// The bytes last copied from the host to one place of a device buffer
// (Frame::memo), under the buffer's allocation number then.
struct UploadMemo {
  uint64_t gen = 0;
  std::vector<unsigned char> bytes;
};

struct Frame {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  // per (buffer, byte offset): the last table uploaded there (upload_table);
  // every buffer takes whole tables at fixed offsets, never overlapping
  std::map<std::pair<const void*, size_t>, UploadMemo> memo;
  std::vector<double> h_grids;  // the allocation-free frame's grid region, on the host
  // the allocation-free frame: every buffer a fixed region of the scratch
  // array; device fields and identity lattices only (their grids go to
  // `grids`)
  bool noalloc = false;
  DevBuf grids;
  std::vector<std::unique_ptr<DevBuf>> uploads;
  DevBuf cnt, tiles, tjobs, rjobs, hdr, jimg;
  DevBuf key, id, rgb, depth, fpix, fdep, frgb, frags, vntmp;
  uint32_t frag_cap = 0;
  HostBuf hrgb, hstage;
  // host sources of asynchronous uploads stay alive until the frame ends
  std::vector<TessJob> h_tjobs;
  std::vector<ObjMeta> h_meta;
  std::vector<uint32_t> obj_job, obj_q;  // field pass and threshold slot of each object
  std::vector<uint32_t> h_jimg;          // images of each field pass (offsets, then indices)
  std::vector<RasterJob<double>> h_rjobs64;
  std::vector<RasterJob<float>> h_rjobs32;
  uint32_t h_nline = 0;
  Frags h_frags;
  contour3d_amd_mesh mesh;  // the frame's triangles, kept between frames
  std::vector<Geom*> geoms_used;  // lattice tables the frame in flight reads
  TessLayout tl;
  // the captured launch sequence of the last frame layout, replayed while
  // the layout (sizes, buffers, capacities) stays the same
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gsig;
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gsig.clear();
  }
  Prof prof;
  bool fp64 = true;         // form of the frame in flight
  FrameHdr last{};          // header of the last finished frame
  bool unread = false;      // the frame in flight was planned with no read back of its totals
  template <typename R>
  std::vector<RasterJob<R>>& rjobs_host();
  hipError_t create() {
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    return e;
  }
  void release() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& u : uploads) u->release();
    uploads.clear();
    DevBuf* bufs[] = {&cnt,  &tiles, &tjobs, &rjobs, &hdr,  &jimg, &key,  &id,   &rgb,
                      &depth, &fpix, &fdep,  &frgb, &frags, &vntmp, &grids};
    for (auto* b : bufs) b->release();
    mesh.release();
    if (!noalloc) frag_cap = 0;
    hrgb.release();
    if (hstage.p) (void)hipHostFree(hstage.p);
    hstage.p = nullptr;
    hstage.bytes = 0;
    hstage.fixed = false;
    for (auto e : prof.pool) (void)hipEventDestroy(e);
    prof.pool.clear();
    prof.pending.clear();
    drop_graph();
    if (done) (void)hipEventDestroy(done);
    done = nullptr;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
  }
};

template <>
std::vector<RasterJob<double>>& Frame::rjobs_host<double>() {
  return h_rjobs64;
}
template <>
std::vector<RasterJob<float>>& Frame::rjobs_host<float>() {
  return h_rjobs32;
}

struct Server;

struct Lib {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int precision = CONTOUR3D_AMD_FP64;
  NbrTable nbr{};
  uint64_t stamp = 0;
  std::vector<std::unique_ptr<Geom>> geoms;
  // workgroups of the tessellation and raster kernels (0: built-in choice)
  uint32_t tess_cap = 0;
  uint32_t raster_cap = 0;
  // tessellation brick shape: rows of cells per tile (0: as many as fit in
  // kTileCells) and planes per work item (0: kDefaultZPlanes)
  uint32_t tess_rows = 0;
  uint32_t tess_zc = 0;
  bool graphs = false;  // replay captured launch sequences of repeated frame layouts
  // Tessellation launches are persistent: by default 128 workgroups per CU,
  // each walking several tiles, so that its triangle queue carries items
  // across tiles and drains part-full once, at its end (measured at 512^3:
  // the fused stage 5.6 -> 5.3 ms against one workgroup per tile)
  uint32_t tess_grid(uint32_t nitems) const {
    const uint32_t cap = tess_cap ? tess_cap : uint32_t(persistent_blocks()) * 16u;
    const uint32_t g = cap < nitems ? cap : nitems;
    return g < 8 ? 8 : g;
  }
  Frame main;
  Frame scratch_frame;  // the allocation-free entry point's frame
  Server* server = nullptr;
};

Lib& lib() {
  static Lib L;
  return L;
}

// Limits of the allocation-free frame (the regions of the scratch array).
constexpr uint64_t kScratchTiles = uint64_t(1) << 20;      // tile-count entries of all objects
constexpr uint32_t kScratchObjs = 64;                      // contour objects
constexpr uint64_t kScratchPixels = uint64_t(1) << 21;     // z-buffer entries (1920x1080 fits)
constexpr uint32_t kScratchFrags = 1u << 19;               // fragment list entries
constexpr uint64_t kScratchLineFrags = uint64_t(1) << 18;  // line fragments
constexpr uint64_t kScratchTriangles = uint64_t(1) << 24;  // cell store entries
constexpr uint64_t kScratchGridValues = uint64_t(1) << 16; // lattice coordinates of all objects

// per device: the scratch array's address there (taken at set-up)
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

// Takes the scratch array's address on the current device and cuts it into
// the buffers of the allocation-free frame F.
int carve_scratch(Frame& F, int device) {
  if (device < 0 || device >= kMaxDevices) {
    log_error("device %d: no scratch array slot", device);
    return 1;
  }
  void* base = nullptr;
  const hipError_t e = hipGetSymbolAddress(&base, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !base) {
    (void)hipGetLastError();
    g_scratch[device] = nullptr;  // the allocation-free entry point reports it; nothing else needs the array
    return 1;
  }
  g_scratch[device] = base;
  size_t off = 0;
  auto region = [&](DevBuf& b, size_t bytes) {
    b.p = static_cast<unsigned char*>(base) + off;
    b.bytes = bytes;
    b.fixed = true;
    b.gen = next_buf_gen();
    off += (bytes + 255) / 256 * 256;
  };
  contour3d_amd_mesh& M = F.mesh;
  region(F.cnt, (kScratchTiles + 1) * sizeof(uint32_t));
  region(M.bstart, (kScratchTiles + 1) * sizeof(uint64_t));
  region(F.tiles, ((kScratchTiles + kScanTile - 1) / kScanTile + 1) * sizeof(uint64_t));
  region(F.tjobs, (kScratchObjs + 1) * sizeof(TessJob));
  region(M.meta, (kScratchObjs + 1) * sizeof(ObjMeta));
  region(F.hdr, sizeof(FrameHdr));
  region(F.jimg, (2 * kScratchObjs + 2) * sizeof(uint32_t));
  region(F.rjobs, sizeof(RasterJob<double>));
  region(F.key, kScratchPixels * sizeof(uint64_t));
  region(F.id, kScratchPixels * sizeof(uint64_t));
  region(F.fpix, (kScratchLineFrags + 1) * sizeof(uint32_t));
  region(F.fdep, (kScratchLineFrags + 1) * sizeof(double));
  region(F.frgb, 3 * (kScratchLineFrags + 1));
  region(F.frags, size_t(kScratchFrags) * sizeof(Frag64));
  region(M.cell, kScratchTriangles * sizeof(uint32_t));
  region(F.grids, kScratchGridValues * sizeof(double));
  region(F.rgb, 0);
  region(F.depth, 0);
  region(M.pos, 0);
  region(M.aux, 0);
  if (off > kScratchBytes) {
    log_error("scratch layout of %zu bytes exceeds the %zu-byte array", off, kScratchBytes);
    return 1;
  }
  M.cap = kScratchTriangles;
  F.frag_cap = kScratchFrags;
  F.noalloc = true;
  return 0;
}

int ensure_init(Lib& L, int device) {
  if (L.ready) return 0;
  int ndev = 0;
  C3D_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    log_error("device %d not available (%d devices)", device, ndev);
    return 1;
  }
  C3D_HIP(hipSetDevice(device));
  C3D_HIP(L.main.create());
  L.nbr = build_nbr_table();
  C3D_HIP(upload_tables(L.nbr));
  // set-up of the allocation-free entry point: its stream, pinned header
  // staging and the scratch array's regions on this device
  Frame& S = L.scratch_frame;
  C3D_HIP(S.create());
  C3D_HIP(S.hstage.ensure(sizeof(FrameHdr) + (kScratchObjs + 1) * sizeof(ObjMeta)));
  S.hstage.fixed = true;
  (void)carve_scratch(S, device);
  L.device = device;
  L.ready = true;
  return 0;
}

// ----------------------------------------------------------------------
// profiling

int prof_begin(Frame& F, int stage) {
  Prof& P = F.prof;
  if (!P.on) return 0;
  if (P.pool.size() < 2) {
    for (int q = 0; q < 8; q++) {
      hipEvent_t e;
      C3D_HIP(hipEventCreate(&e));
      P.pool.push_back(e);
    }
  }
  P.cur_a = P.pool.back();
  P.pool.pop_back();
  P.cur_stage = stage;
  C3D_HIP(hipEventRecord(P.cur_a, F.st));
  return 0;
}

int prof_end(Frame& F) {
  Prof& P = F.prof;
  if (!P.on || P.cur_stage < 0) return 0;
  hipEvent_t b = P.pool.back();
  P.pool.pop_back();
  C3D_HIP(hipEventRecord(b, F.st));
  P.pending.push_back({P.cur_stage, P.cur_a, b});
  P.cur_stage = -1;
  return 0;
}

// a stage timed within the open one (the fused kernel inside the raster stage)
int prof_inner_begin(Frame& F, int stage) {
  Prof& P = F.prof;
  if (!P.on) return 0;
  if (P.pool.size() < 2) {
    for (int q = 0; q < 8; q++) {
      hipEvent_t e;
      C3D_HIP(hipEventCreate(&e));
      P.pool.push_back(e);
    }
  }
  P.in_a = P.pool.back();
  P.pool.pop_back();
  P.in_stage = stage;
  C3D_HIP(hipEventRecord(P.in_a, F.st));
  return 0;
}

int prof_inner_end(Frame& F) {
  Prof& P = F.prof;
  if (!P.on || P.in_stage < 0) return 0;
  hipEvent_t b = P.pool.back();
  P.pool.pop_back();
  C3D_HIP(hipEventRecord(b, F.st));
  P.pending.push_back({P.in_stage, P.in_a, b});
  P.in_stage = -1;
  return 0;
}

// after a stream synchronisation
int prof_collect(Frame& F) {
  Prof& P = F.prof;
  for (auto& q : P.pending) {
    float ms = 0;
    C3D_HIP(hipEventElapsedTime(&ms, q.a, q.b));
    P.ms[q.stage] += ms;
    P.calls[q.stage] += 1;
    P.pool.push_back(q.a);
    P.pool.push_back(q.b);
  }
  P.pending.clear();
  return 0;
}

// ----------------------------------------------------------------------
// pointers, fields and geometry

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t a;
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// FNV-1a over 8-byte words (then the remaining bytes): a filter for the
// geometry cache, whose hits are confirmed by comparing the values
uint64_t fnv(const void* data, size_t bytes, uint64_t h) {
  const unsigned char* c = static_cast<const unsigned char*>(data);
  size_t q = 0;
  for (; q + 8 <= bytes; q += 8) {
    uint64_t w;
    std::memcpy(&w, c + q, 8);
    h ^= w;
    h *= 1099511628211ull;
  }
  for (; q < bytes; q++) {
    h ^= c[q];
    h *= 1099511628211ull;
  }
  return h;
}

using ConvFn = contour3d_vector_t (*)(const contour3d_vector_t);

// Device coordinates of an object's lattice.  The converter maps each
// lattice point (grids[0][i], grids[1][j], grids[2][k]) to Cartesian
// (triangulate.c:165-170); it is evaluated once per point and the result
// cached for as long as the converter, grid pointers and grid values stay
// the same.  A map that returns every point unchanged needs no table.
int get_geometry(Lib& L, const contour3d_contour_obj_t& o, Geom** out) {
  const size_t nx = o.glsizes[0], ny = o.glsizes[1], nz = o.glsizes[2];
  std::vector<double> hg[3];
  const double* gp[3];
  for (int d = 0; d < 3; d++) {
    if (is_device_ptr(o.grids[d])) {
      hg[d].resize(o.glsizes[d]);
      C3D_HIP(hipMemcpy(hg[d].data(), o.grids[d], o.glsizes[d] * sizeof(double), hipMemcpyDeviceToHost));
      gp[d] = hg[d].data();
    } else {
      gp[d] = o.grids[d];
    }
  }
  uint64_t h = 1469598103934665603ull;
  for (int d = 0; d < 3; d++) h = fnv(gp[d], o.glsizes[d] * sizeof(double), h);
  // a geometry is its converter and grid values: callers that rebuild equal
  // grid arrays for every call still hit the cache
  for (auto& g : L.geoms) {
    if (g->conv == (void*)o.converter && g->hash == h && g->n[0] == nx && g->n[1] == ny && g->n[2] == nz &&
        std::memcmp(g->host.data(), gp[0], nx * sizeof(double)) == 0 &&
        std::memcmp(g->host.data() + nx, gp[1], ny * sizeof(double)) == 0 &&
        std::memcmp(g->host.data() + nx + ny, gp[2], nz * sizeof(double)) == 0) {
      g->stamp = ++L.stamp;
      *out = g.get();
      return 0;
    }
  }
  auto g = std::make_unique<Geom>();
  g->conv = (void*)o.converter;
  for (int d = 0; d < 3; d++) {
    g->g[d] = o.grids[d];
    g->n[d] = o.glsizes[d];
  }
  g->hash = h;
  g->host.assign(gp[0], gp[0] + nx);
  g->host.insert(g->host.end(), gp[1], gp[1] + ny);
  g->host.insert(g->host.end(), gp[2], gp[2] + nz);
  // a geometry that fails half-built returns its device memory
  struct Unbuilt {
    Geom* g;
    ~Unbuilt() {
      if (g) {
        g->grids.release();
        g->table.release();
      }
    }
  } unbuilt{g.get()};
  C3D_HIP(g->grids.ensure((nx + ny + nz) * sizeof(double)));
  C3D_HIP(hipMemcpy(g->grids.as<double>(), gp[0], nx * sizeof(double), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(g->grids.as<double>() + nx, gp[1], ny * sizeof(double), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(g->grids.as<double>() + nx + ny, gp[2], nz * sizeof(double), hipMemcpyHostToDevice));
  g->identity = true;
  if (o.converter != contour3d_amd_identity) {
    const ConvFn conv = o.converter;
    const size_t npts = nx * ny * nz;
    std::vector<double> tab;
    size_t p = 0;
    for (size_t k = 0; k < nz; k++) {
      for (size_t j = 0; j < ny; j++) {
        for (size_t i = 0; i < nx; i++, p++) {
          const contour3d_vector_t in = {gp[0][i], gp[1][j], gp[2][k]};
          const contour3d_vector_t r = conv(in);
          if (g->identity && std::memcmp(&r, &in, sizeof(r)) != 0) {
            g->identity = false;
            tab.resize(3 * npts);
            size_t q = 0;
            for (size_t kk = 0; kk <= k; kk++)
              for (size_t jj = 0; jj < ny; jj++)
                for (size_t ii = 0; ii < nx && q < p; ii++, q++) {
                  tab[q] = gp[0][ii];
                  tab[npts + q] = gp[1][jj];
                  tab[2 * npts + q] = gp[2][kk];
                }
          }
          if (!g->identity) {
            tab[p] = r.x;
            tab[npts + p] = r.y;
            tab[2 * npts + p] = r.z;
          }
        }
      }
    }
    if (!g->identity) {
      C3D_HIP(g->table.ensure(3 * npts * sizeof(double)));
      C3D_HIP(hipMemcpy(g->table.p, tab.data(), 3 * npts * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  unbuilt.g = nullptr;
  g->stamp = ++L.stamp;
  // Keep up to kMaxGeoms geometries and kGeomTableBytes of coordinate
  // tables; evict the least recently used ones that no frame in flight
  // reads.  A grid-only geometry holds a few kilobytes, so a batch of many
  // lattices stays cached: an eviction frees device memory, and a free
  // waits for the device.
  size_t table_bytes = g->table.bytes;
  for (auto& e : L.geoms) table_bytes += e->table.bytes;
  while (L.geoms.size() >= kMaxGeoms || table_bytes > kGeomTableBytes) {
    size_t victim = L.geoms.size();
    for (size_t q = 0; q < L.geoms.size(); q++)
      if (L.geoms[q]->busy.load() == 0 && (victim == L.geoms.size() || L.geoms[q]->stamp < L.geoms[victim]->stamp))
        victim = q;
    if (victim == L.geoms.size()) break;
    table_bytes -= L.geoms[victim]->table.bytes;
    L.geoms[victim]->grids.release();
    L.geoms[victim]->table.release();
    L.geoms.erase(L.geoms.begin() + victim);
  }
  *out = g.get();
  L.geoms.push_back(std::move(g));
  return 0;
}

int validate_obj(const contour3d_contour_obj_t& o, size_t n) {
  for (int d = 0; d < 3; d++) {
    if (o.glsizes[d] < 2) {
      log_error("contour object %zu: glsizes[%d] = %zu, need at least 2 points per axis", n, d, o.glsizes[d]);
      return 1;
    }
    if (!o.grids[d]) {
      log_error("contour object %zu: grids[%d] is NULL", n, d);
      return 1;
    }
  }
  if (!o.array || !o.converter) {
    log_error("contour object %zu: array and converter must not be NULL", n);
    return 1;
  }
  const uint64_t nc = uint64_t(o.glsizes[0] - 1) * uint64_t(o.glsizes[1] - 1) * uint64_t(o.glsizes[2] - 1);
  if (nc >= (uint64_t(1) << 32) - kTileCells || o.glsizes[0] > (1u << 30) || o.glsizes[1] > (1u << 30) ||
      o.glsizes[2] > (1u << 30)) {
    log_error("contour object %zu: %zu x %zu x %zu lattice exceeds 2^32 cells", n, o.glsizes[0], o.glsizes[1],
              o.glsizes[2]);
    return 1;
  }
  return 0;
}

template <typename R>
MeshView<R> view_of(const contour3d_amd_mesh& M) {
  MeshView<R> m;
  m.pos = M.pos.as<MeshPos<R>>();
  m.aux = M.aux.as<MeshAux<R>>();
  m.cap = M.cap;
  m.code = M.code.as<uint8_t>();
  return m;
}

// ----------------------------------------------------------------------
// view and lines (host FP64, exactly as the reference forms them)

using V = contour3d_vector_t;
V hsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V hadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V hscale(double w, V a) { return {w * a.x, w * a.y, w * a.z}; }
double hdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V hcross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V hunit(V a) {
  const double s = 1. / std::sqrt(hdot(a, a));
  return {a.x * s, a.y * s, a.z * s};
}

struct HostView {
  V cam, ctr, lx, ly, nrm, light;
  double nc, d, lxx, lyy;
  size_t W, H;
  contour3d_color_t bg;
};

// src/contour3d/main.c:25-63 and the per-call constants of project.c:34-36
HostView make_view(const V* cam, const V* light, const size_t sizes[2], const V* ctr, const V local[2],
                   const contour3d_color_t* bg) {
  HostView v;
  v.cam = *cam;
  v.ctr = *ctr;
  v.lx = local[0];
  v.ly = local[1];
  v.light = hunit(*light);
  v.nrm = hunit(hcross(local[0], local[1]));
  v.nc = hdot(v.nrm, v.cam);
  v.d = hdot(v.nrm, v.ctr);
  v.lxx = hdot(v.lx, v.lx);
  v.lyy = hdot(v.ly, v.ly);
  v.W = sizes[0];
  v.H = sizes[1];
  v.bg = *bg;
  return v;
}

template <typename R>
ViewDev<R> dev_view(const HostView& h) {
  ViewDev<R> v;
  const V* src[6] = {&h.cam, &h.ctr, &h.lx, &h.ly, &h.nrm, &h.light};
  R* dst[6] = {v.cam, v.ctr, v.lx, v.ly, v.nrm, v.light};
  for (int q = 0; q < 6; q++) {
    dst[q][0] = R(src[q]->x);
    dst[q][1] = R(src[q]->y);
    dst[q][2] = R(src[q]->z);
  }
  v.nc = R(h.nc);
  v.d = R(h.d);
  v.lxx = R(h.lxx);
  v.lyy = R(h.lyy);
  v.fw = R(double(h.W));
  v.fh = R(double(h.H));
  set_pitch(v);
  v.W = int32_t(h.W);
  v.H = int32_t(h.H);
  return v;
}

// project.c:9-67 on the host
bool hproject(const HostView& v, V p, V& q) {
  const V ray = hsub(p, v.cam);
  const double nr = hdot(v.nrm, ray);
  if (-1.e-8 < nr && nr < 1.e-8) return false;
  const double t = (v.d - v.nc) / nr;
  if (t <= 0. || 1. <= t) return false;
  const V delta = {v.cam.x + t * ray.x - v.ctr.x, v.cam.y + t * ray.y - v.ctr.y, v.cam.z + t * ray.z - v.ctr.z};
  q.x = hdot(delta, v.lx) / v.lxx;
  q.y = hdot(delta, v.ly) / v.lyy;
  q.z = nr;
  return true;
}

// Line objects (line.c:26-170) become a list of (pixel, depth) fragments in
// the reference's drawing order; the z-test against the contours happens on
// the device.  Coverage uses libm pow as the reference does, whose rounding
// differs from a plain square in about 1 of 1200 arguments.
int line_fragments(const HostView& v, size_t nlines, const contour3d_line_obj_t* lines, Frags& F) {
  const long W = long(v.W), H = long(v.H);
  for (size_t ln = 0; ln < nlines; ln++) {
    const contour3d_line_obj_t& L = lines[ln];
    if (L.nitems < 2) {
      log_error("line object %zu: give more than two points to draw a line, received %zu", ln, L.nitems);
      return 1;
    }
    if (!L.converter) {
      log_error("line object %zu: converter is NULL", ln);
      return 1;
    }
    const V head = L.edges[0], tail = L.edges[1];
    const V delta = hscale(1. / double(L.nitems - 1), hsub(tail, head));
    for (size_t s = 0; s + 1 < L.nitems; s++) {
      const V a0 = L.converter(hadd(head, hscale(double(s + 0), delta)));
      const V b0 = L.converter(hadd(head, hscale(double(s + 1), delta)));
      V a, b;
      if (!hproject(v, a0, a) || !hproject(v, b0, b)) continue;
      a.x = (0.5 + a.x) * double(v.W);
      a.y = (0.5 + a.y) * double(v.H);
      b.x = (0.5 + b.x) * double(v.W);
      b.y = (0.5 + b.y) * double(v.H);
      const long xmin = long(std::fmin(a.x, b.x) - 1), xmax = long(std::fmax(a.x, b.x) + 1);
      const long ymin = long(std::fmin(a.y, b.y) - 1), ymax = long(std::fmax(a.y, b.y) + 1);
      auto clampi = [](long lo, long x, long hi) { return x < lo ? lo : (x > hi ? hi : x); };
      const long i0 = clampi(0, xmin, W - 1), i1 = clampi(0, xmax, W - 1);
      const long j0 = clampi(0, ymin, H - 1), j1 = clampi(0, ymax, H - 1);
      const double la = b.y - a.y;
      const double lb = -b.x + a.x;
      const double lc = (b.x - a.x) * a.y - (b.y - a.y) * a.x;
      for (long j = j0; j <= j1; j++) {
        for (long i = i0; i <= i1; i++) {
          const double di = double(i), dj = double(j);
          const double lim = std::pow(0.5 * L.width, 2.);
          const double dist = std::pow(la * di + lb * dj + lc, 2.) / (la * la + lb * lb);
          if (lim < dist) continue;
          const double prm = -1. * ((b.x - a.x) * (a.x - di) + (b.y - a.y) * (a.y - dj)) /
                             (std::pow(b.x - a.x, 2.) + std::pow(b.y - a.y, 2.));
          const double dep = 1. / (prm / b.z + (1. - prm) / a.z);
          F.pix.push_back(uint32_t(j * W + i));
          F.depth.push_back(dep);
          F.rgb.push_back(L.color.r);
          F.rgb.push_back(L.color.g);
          F.rgb.push_back(L.color.b);
        }
      }
    }
  }
  return 0;
}

// ----------------------------------------------------------------------
// frames
//
// A frame is a list of images.  Every contour object of the frame is
// tessellated by one count, scan, plan and emit sequence over a device table
// of field passes; every image is then rasterised and resolved by one
// launch per stage over a device table of images.  An image whose object
// list equals an earlier image's (same fields, lattices, converters,
// thresholds and colours) reuses that image's triangles.

struct Image {
  HostView hv;
  size_t nobj = 0;
  const contour3d_contour_obj_t* objs = nullptr;
  size_t nlines = 0;
  const contour3d_line_obj_t* lines = nullptr;
  uint8_t* rgb = nullptr;  // caller's buffer
  double* depth = nullptr; // caller's buffer or null
  bool rgb_dev = false, dep_dev = false;
  uint32_t obj0 = 0, obj1 = 0;
  uint64_t px0 = 0;
};

bool same_obj(const contour3d_contour_obj_t& a, const contour3d_contour_obj_t& b) {
  return a.array == b.array && a.converter == b.converter && a.pencil == b.pencil &&
         std::memcmp(a.glsizes, b.glsizes, sizeof(a.glsizes)) == 0 &&
         std::memcmp(a.grids, b.grids, sizeof(a.grids)) == 0 &&
         std::memcmp(&a.threshold, &b.threshold, sizeof(double)) == 0 && a.color.r == b.color.r &&
         a.color.g == b.color.g && a.color.b == b.color.b;
}

// A frame keeps only the cell index of each triangle (the fused raster
// draws triangles from registers): 1.25x the triangles seen.
// Copies n bytes of host memory to device buffer b at byte offset off on
// F's stream, unless the same bytes were the last ones copied there and b
// has not been reallocated since: a repeated frame then sends only the
// tables that changed (each small copy from pageable memory costs about
// 12 us of host time, and the device waits for it).  Kernels may rewrite
// parts of an uploaded table (k_plan: objects' bases and counts, images'
// triangle ranges) only where every frame that reads them rewrites them.
int upload_table(Frame& F, DevBuf& b, size_t off, const void* src, size_t n) {
  if (n == 0) return 0;
  UploadMemo& m = F.memo[{static_cast<const void*>(&b), off}];
  if (b.gen != 0 && m.gen == b.gen && m.bytes.size() == n && std::memcmp(m.bytes.data(), src, n) == 0) return 0;
  m.gen = 0;
  C3D_HIP(hipMemcpyAsync(b.as<unsigned char>() + off, src, n, hipMemcpyHostToDevice, F.st));
  m.bytes.assign(static_cast<const unsigned char*>(src), static_cast<const unsigned char*>(src) + n);
  m.gen = b.gen;
  return 0;
}

// forgets what upload_table last sent to (b, off): something else wrote it
void forget_table(Frame& F, const DevBuf& b, size_t off) { F.memo.erase({static_cast<const void*>(&b), off}); }

int ensure_cells(contour3d_amd_mesh& M) {
  if (M.cap >= M.total && M.cap != 0 && M.cell.p) return 0;
  const uint64_t cap = M.total + M.total / 4 + 1024;
  M.cap = 0;
  C3D_HIP(M.cell.ensure(cap * sizeof(uint32_t), true));
  M.cap = cap;
  return 0;
}

template <typename R>
int ensure_capacity(contour3d_amd_mesh& M) {
  C3D_HIP(M.code.ensure(M.code_bytes, true));
  if (M.cap >= M.total && M.cap != 0 && M.precision_built == M.precision) return 0;
  M.precision_built = M.precision;
  const uint64_t cap = M.total + M.total / 4 + 1024;
  M.pos.release();
  M.aux.release();
  M.cap = 0;
  // the capacity already holds a quarter of headroom
  C3D_HIP(M.pos.ensure(cap * sizeof(MeshPos<R>), true));
  C3D_HIP(M.aux.ensure(cap * sizeof(MeshAux<R>), true));
  M.cap = cap;
  return 0;
}

// Host half of a frame's tessellation: validates the objects, resolves
// fields (host fields are uploaded) and lattices, groups thresholds of a
// shared field into field passes, sizes the buffers and enqueues the uploads
// of the pass and object tables.  The layout goes to F.tl.
template <typename R>
int tess_prepare(Lib& L, Frame& F, contour3d_amd_mesh& M, const std::vector<const contour3d_contour_obj_t*>& objs) {
  const size_t n = objs.size();
  if (F.noalloc && n > F.tjobs.bytes / sizeof(TessJob) - 1) {
    log_error("the allocation-free call takes at most %zu contour objects", F.tjobs.bytes / sizeof(TessJob) - 1);
    return 1;
  }
  size_t grid_used = 0;  // lattice coordinates in the allocation-free frame's grid region
  // and where each object's grid goes there: (first value, source, values)
  struct GridPiece {
    size_t at;
    const double* src;
    size_t n;
  };
  std::vector<GridPiece> grid_pieces;
  std::vector<FieldDev> fd(n);
  std::vector<int> group(n, -1);
  std::vector<uint32_t> boff(n + 1, 0);
  std::vector<const double*> upl_src;
  std::vector<void*> upl_dst;
  // per object: whether its array is device memory, and its lattice; an
  // object with the array, or the converter, sizes and grid pointers, of an
  // earlier object of this call takes that object's (the frame's objects
  // usually share one field and lattice)
  std::vector<char> on_dev(n, 0);
  std::vector<Geom*> geo(n, nullptr);
  for (size_t o = 0; o < n; o++) {
    const contour3d_contour_obj_t& ob = *objs[o];
    if (validate_obj(ob, o)) return 1;
    int same_array = -1, same_lattice = -1;
    for (size_t q = 0; q < o; q++) {
      const contour3d_contour_obj_t& oq = *objs[q];
      if (same_array < 0 && oq.array == ob.array) same_array = int(q);
      if (same_lattice < 0 && oq.converter == ob.converter &&
          std::memcmp(oq.glsizes, ob.glsizes, sizeof(ob.glsizes)) == 0 &&
          std::memcmp(oq.grids, ob.grids, sizeof(ob.grids)) == 0)
        same_lattice = int(q);
    }
    on_dev[o] = same_array >= 0 ? on_dev[same_array] : char(is_device_ptr(ob.array));
    FieldDev& f = fd[o];
    f.nx = uint32_t(ob.glsizes[0]);
    f.ny = uint32_t(ob.glsizes[1]);
    f.nz = uint32_t(ob.glsizes[2]);
    f.ncx = f.nx - 1;
    f.ncy = f.ny - 1;
    f.ncz = f.nz - 1;
    f.ncells = f.ncx * f.ncy * f.ncz;
    f.div_ncx = make_fastdiv(f.ncx);
    f.div_ncy = make_fastdiv(f.ncy);
    f.tg = make_tile_geo(f.ncx, f.ncy, f.ncz, L.tess_rows, L.tess_zc);
    f.ntiles = f.ncz * f.tg.cols;  // at most the cell count
    f.nitems = f.tg.cols * f.tg.nzc;
    if (F.noalloc) {
      // the allocation-free frame reads device fields in place and the
      // identity map of grids copied into its scratch region
      if (!on_dev[o] || ob.converter != contour3d_amd_identity) {
        log_error("contour object %zu: the allocation-free call takes device fields and the "
                  "contour3d_amd_identity converter", o);
        return 1;
      }
      f.f = ob.array;
      const double* reuse = nullptr;
      for (size_t q = 0; q < o && !reuse; q++)
        if (std::memcmp(objs[q]->glsizes, ob.glsizes, sizeof(ob.glsizes)) == 0 &&
            std::memcmp(objs[q]->grids, ob.grids, sizeof(ob.grids)) == 0)
          reuse = fd[q].gx;
      if (!reuse) {
        const size_t need = ob.glsizes[0] + ob.glsizes[1] + ob.glsizes[2];
        if (grid_used + need > F.grids.bytes / sizeof(double)) {
          log_error("the allocation-free call holds at most %zu lattice coordinates (grids of all objects)",
                    F.grids.bytes / sizeof(double));
          return 1;
        }
        double* gd = F.grids.as<double>() + grid_used;
        size_t go = 0;
        for (int d = 0; d < 3; d++) {
          grid_pieces.push_back({grid_used + go, ob.grids[d], size_t(ob.glsizes[d])});
          go += ob.glsizes[d];
        }
        grid_used += need;
        reuse = gd;
      }
      f.gx = reuse;
      f.gy = f.gx + f.nx;
      f.gz = f.gy + f.ny;
      f.cx = f.cy = f.cz = nullptr;
    } else if (on_dev[o]) {
      f.f = ob.array;
    } else {
      const size_t bytes = size_t(f.nx) * f.ny * f.nz * sizeof(double);
      void* dst = nullptr;
      for (size_t q = 0; q < upl_src.size(); q++)
        if (upl_src[q] == ob.array) dst = upl_dst[q];
      if (!dst) {
        if (F.uploads.size() <= upl_src.size()) F.uploads.push_back(std::make_unique<DevBuf>());
        DevBuf& b = *F.uploads[upl_src.size()];
        C3D_HIP(b.ensure(bytes));
        C3D_HIP(hipMemcpyAsync(b.p, ob.array, bytes, hipMemcpyHostToDevice, F.st));
        upl_src.push_back(ob.array);
        upl_dst.push_back(b.p);
        dst = b.p;
      }
      f.f = static_cast<const double*>(dst);
    }
    if (!F.noalloc) {
      Geom* g = same_lattice >= 0 ? geo[same_lattice] : nullptr;
      if (!g && get_geometry(L, ob, &g)) return 1;
      geo[o] = g;
      g->busy++;
      F.geoms_used.push_back(g);
      f.gx = g->grids.as<double>();
      f.gy = f.gx + f.nx;
      f.gz = f.gy + f.ny;
      if (g->identity) {
        f.cx = f.cy = f.cz = nullptr;
      } else {
        const size_t npts = size_t(f.nx) * f.ny * f.nz;
        f.cx = g->table.as<double>();
        f.cy = f.cx + npts;
        f.cz = f.cy + npts;
      }
    }
    if (uint64_t(boff[o]) + f.ntiles >= (uint64_t(1) << 31)) {
      log_error("frame exceeds 2^31 tessellation tiles");
      return 1;
    }
    boff[o + 1] = boff[o] + f.ntiles;
    for (size_t q = 0; q < o && group[o] < 0; q++) {
      if (fd[q].f == f.f && fd[q].nx == f.nx && fd[q].ny == f.ny && fd[q].nz == f.nz && fd[q].gx == f.gx &&
          fd[q].cx == f.cx)
        group[o] = group[q];
    }
    if (group[o] < 0) group[o] = int(o);
  }
  const uint64_t nbt = boff[n];
  if (F.noalloc && nbt + 1 > F.cnt.bytes / sizeof(uint32_t)) {
    log_error("the allocation-free call holds at most %zu tessellation tiles; the objects need %llu",
              F.cnt.bytes / sizeof(uint32_t) - 1, (unsigned long long)nbt);
    return 1;
  }
  M.precision = sizeof(R) == 8 ? CONTOUR3D_AMD_FP64 : CONTOUR3D_AMD_FP32;
  M.base.assign(n, 0);
  M.count.assign(n, 0);
  M.tiles.resize(n);
  M.geo.resize(n);
  for (size_t o = 0; o < n; o++) {
    M.tiles[o] = fd[o].ntiles;
    M.geo[o] = fd[o].tg;
  }
  M.color.resize(n);
  for (size_t o = 0; o < n; o++) M.color[o] = objs[o]->color;
  M.ntiles = nbt;

  // field passes: thresholds of objects sharing a field are classified in
  // one read of it, up to kMaxThr at a time
  std::vector<TessJob>& tj = F.h_tjobs;
  tj.clear();
  F.obj_job.assign(n, 0);
  F.obj_q.assign(n, 0);
  uint64_t jt = 0, jx = 0, codes = 0;
  std::vector<uint64_t> code_off(n, 0);
  for (size_t o = 0; o < n; o++) {
    if (group[o] != int(o)) continue;
    TessJob job{};
    job.f = fd[o];
    job.code_off = codes;
    for (size_t q = o; q < n; q++) {
      if (group[q] != int(o)) continue;
      code_off[q] = codes;
      codes += fd[o].ncells;
      job.ts.thr[job.ts.n] = objs[q]->threshold;
      job.ts.boff[job.ts.n] = boff[q];
      job.ts.obj[job.ts.n] = uint32_t(q);
      F.obj_job[q] = uint32_t(tj.size());
      F.obj_q[q] = uint32_t(job.ts.n);
      job.ts.n++;
      if (job.ts.n == kMaxThr) {
        job.item0 = uint32_t(jt);
        job.xoff = uint32_t(jx);
        jt += fd[o].nitems;
        jx += (fd[o].nitems + 7) / 8;
        tj.push_back(job);
        job.ts = ThrSet{};
        job.code_off = codes;
      }
    }
    if (job.ts.n) {
      job.item0 = uint32_t(jt);
      job.xoff = uint32_t(jx);
      jt += fd[o].nitems;
      jx += (fd[o].nitems + 7) / 8;
      tj.push_back(job);
    }
  }
  if (jt >= (uint64_t(1) << 31)) {
    log_error("frame exceeds 2^31 tessellation work items");
    return 1;
  }
  const uint32_t ntj = uint32_t(jt);
  M.code_bytes = codes;  // a stored mesh keeps every cell's corner code
  M.starts_ready = false;

  C3D_HIP(F.cnt.ensure((nbt + 1) * sizeof(uint32_t)));
  C3D_HIP(M.bstart.ensure((nbt + 1) * sizeof(uint64_t)));
  C3D_HIP(F.tiles.ensure(((nbt + kScanTile - 1) / kScanTile + 1) * sizeof(uint64_t)));
  C3D_HIP(F.tjobs.ensure((tj.size() + 1) * sizeof(TessJob)));
  C3D_HIP(M.meta.ensure((n + 1) * sizeof(ObjMeta)));
  C3D_HIP(F.hdr.ensure(sizeof(FrameHdr)));
  C3D_HIP(F.hstage.ensure(sizeof(FrameHdr) + (n + 1) * sizeof(ObjMeta)));
  std::vector<ObjMeta>& hm = F.h_meta;
  hm.assign(n, ObjMeta{});
  for (size_t o = 0; o < n; o++) {
    hm[o].bstart = M.bstart.as<uint64_t>() + boff[o];
    hm[o].ncx = fd[o].ncx;
    hm[o].ncy = fd[o].ncy;
    hm[o].ncz = fd[o].ncz;
    hm[o].div_ncx = fd[o].div_ncx;
    hm[o].div_ncy = fd[o].div_ncy;
    hm[o].tg = fd[o].tg;
    hm[o].tile_off = boff[o];
    hm[o].job = F.obj_job[o];
    hm[o].q = F.obj_q[o];
    hm[o].code_off = code_off[o];
    hm[o].rgb[0] = objs[o]->color.r;
    hm[o].rgb[1] = objs[o]->color.g;
    hm[o].rgb[2] = objs[o]->color.b;
  }
  if (!grid_pieces.empty()) {
    // the grid region as one table (its pieces change places with the
    // objects' lattices), unless a grid lives on the device
    bool dev = false;
    for (const GridPiece& g : grid_pieces) dev = dev || is_device_ptr(g.src);
    if (dev) {
      forget_table(F, F.grids, 0);
      for (const GridPiece& g : grid_pieces)
        C3D_HIP(hipMemcpyAsync(F.grids.as<double>() + g.at, g.src, g.n * sizeof(double), hipMemcpyDefault, F.st));
    } else {
      F.h_grids.resize(grid_used);
      for (const GridPiece& g : grid_pieces) std::memcpy(F.h_grids.data() + g.at, g.src, g.n * sizeof(double));
      if (upload_table(F, F.grids, 0, F.h_grids.data(), grid_used * sizeof(double))) return 1;
    }
  }
  if (n) {
    if (upload_table(F, M.meta, 0, hm.data(), n * sizeof(ObjMeta))) return 1;
    if (upload_table(F, F.tjobs, 0, tj.data(), tj.size() * sizeof(TessJob))) return 1;
  }
  F.tl.nbt = nbt;
  F.tl.ntj = ntj;
  F.tl.njobs = int(tj.size());
  F.tl.nobj = int(n);
  return 0;
}

// count, scan and plan; the plan flags a frame whose triangles exceed cap
template <typename R>
int tess_head(Lib& L, Frame& F, contour3d_amd_mesh& M, RasterJob<R>* rjobs, int njobs, uint64_t cap) {
  const TessLayout& tl = F.tl;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_COUNT)) return 1;
  C3D_HIP(launch_count(F.tjobs.as<TessJob>(), tl.njobs, tl.ntj, F.cnt.as<uint32_t>(), L.tess_grid(tl.ntj), F.st));
  if (prof_end(F)) return 1;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_SCAN)) return 1;
  if (tl.nbt)
    C3D_HIP(launch_scan(F.cnt.as<uint32_t>(), M.bstart.as<uint64_t>(), tl.nbt, F.tiles.as<uint64_t>(), F.st));
  C3D_HIP(launch_plan<R>(M.bstart.as<uint64_t>(), tl.nbt, M.meta.as<ObjMeta>(), tl.nobj, rjobs, njobs,
                         F.hdr.as<FrameHdr>(), cap, F.st));
  if (prof_end(F)) return 1;
  return 0;
}

template <typename R>
int tess_emit(Lib& L, Frame& F, contour3d_amd_mesh& M) {
  const TessLayout& tl = F.tl;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_EMIT)) return 1;
  // the store-bound emit keeps one workgroup per tile unless capped
  C3D_HIP(launch_emit<R>(F.tjobs.as<TessJob>(), tl.njobs, tl.ntj, M.bstart.as<uint64_t>(), view_of<R>(M),
                         F.hdr.as<FrameHdr>(), L.tess_cap ? L.tess_grid(tl.ntj) : tl.ntj, F.st));
  if (prof_end(F)) return 1;
  return 0;
}

// Reads the plan's totals back (one synchronisation) and sizes the mesh
// (only its cell indices for a frame).
template <typename R>
int tess_readback(Frame& F, contour3d_amd_mesh& M, FrameHdr& hh, uint64_t nline, bool cells_only = false) {
  const size_t n = size_t(F.tl.nobj);
  unsigned char* hs = static_cast<unsigned char*>(F.hstage.p);
  C3D_HIP(hipMemcpyAsync(hs, F.hdr.p, sizeof(FrameHdr), hipMemcpyDeviceToHost, F.st));
  if (n)
    C3D_HIP(hipMemcpyAsync(hs + sizeof(FrameHdr), M.meta.p, n * sizeof(ObjMeta), hipMemcpyDeviceToHost, F.st));
  C3D_HIP(hipStreamSynchronize(F.st));
  std::memcpy(&hh, hs, sizeof(FrameHdr));
  for (size_t o = 0; o < n; o++) {
    ObjMeta om;
    std::memcpy(&om, hs + sizeof(FrameHdr) + o * sizeof(ObjMeta), sizeof(ObjMeta));
    M.base[o] = om.base;
    M.count[o] = om.count;
  }
  M.total = hh.ntri;
  F.prof.ntri += M.total;
  if (!frame_ids_fit(M.total, nline)) {
    log_error("%llu triangles and %llu line fragments exceed the 2^32 ids of one frame",
              (unsigned long long)M.total, (unsigned long long)nline);
    return 1;
  }
  return cells_only ? ensure_cells(M) : ensure_capacity<R>(M);
}

// Count, scan, plan and emit every object of `objs` into M.  `rjobs`
// (device, njobs entries) receive their triangle ranges and work offsets
// from the plan kernel.  Synchronises once, after the plan, to size the
// mesh; on return the emission is enqueued.
template <typename R>
int tessellate(Lib& L, Frame& F, contour3d_amd_mesh& M, const std::vector<const contour3d_contour_obj_t*>& objs,
               RasterJob<R>* rjobs, int njobs, FrameHdr& hh) {
  if (tess_prepare<R>(L, F, M, objs)) return 1;
  // a mesh has no line fragments (the plan checks the frame's ids)
  forget_table(F, F.hdr, offsetof(FrameHdr, nline));
  C3D_HIP(hipMemsetAsync(F.hdr.as<unsigned char>() + offsetof(FrameHdr, nline), 0, sizeof(uint32_t), F.st));
  if (tess_head<R>(L, F, M, rjobs, njobs, ~uint64_t(0))) return 1;
  if (tess_readback<R>(F, M, hh, 0)) return 1;
  return tess_emit<R>(L, F, M);
}

// Line fragments of every image, appended to F.h_frags with frame pixel
// indices.
int frame_lines(Frame& F, const std::vector<Image>& imgs) {
  Frags& LF = F.h_frags;
  LF.pix.clear();
  LF.depth.clear();
  LF.rgb.clear();
  for (const Image& im : imgs) {
    if (im.nlines && !im.lines) {
      log_error("line objects pointer is NULL");
      return 1;
    }
    const size_t f0 = LF.pix.size();
    if (line_fragments(im.hv, im.nlines, im.lines, LF)) return 1;
    for (size_t q = f0; q < LF.pix.size(); q++) LF.pix[q] += uint32_t(im.px0);
  }
  if (LF.pix.size() >= (size_t(1) << 31)) {
    log_error("too many line fragments in one frame");
    return 1;
  }
  return 0;
}

// Uploads the frame's line fragments (and their count into the header) and
// sizes the z-buffer and the FP64 fragment list.
template <typename R>
int raster_prepare(Frame& F, uint64_t npx) {
  const bool fp64 = sizeof(R) == 8;
  const Frags& LF = F.h_frags;
  const uint32_t nf = uint32_t(LF.pix.size());
  if (F.noalloc && (nf + 1 > F.fpix.bytes / sizeof(uint32_t) || npx > F.key.bytes / sizeof(uint64_t))) {
    log_error("the allocation-free call holds at most %zu line fragments and %zu pixels (%u and %llu asked)",
              F.fpix.bytes / sizeof(uint32_t) - 1, F.key.bytes / sizeof(uint64_t), nf, (unsigned long long)npx);
    return 1;
  }
  C3D_HIP(F.key.ensure(npx * sizeof(uint64_t)));
  // FP64: 64-bit ids for fused frames (32-bit for stored meshes)
  if (fp64) C3D_HIP(F.id.ensure(npx * sizeof(uint64_t)));
  C3D_HIP(F.fpix.ensure((nf + 1) * sizeof(uint32_t)));
  C3D_HIP(F.fdep.ensure((nf + 1) * sizeof(double)));
  C3D_HIP(F.frgb.ensure(3 * size_t(nf + 1)));
  if (nf) {
    C3D_HIP(hipMemcpyAsync(F.fpix.p, LF.pix.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice, F.st));
    C3D_HIP(hipMemcpyAsync(F.fdep.p, LF.depth.data(), nf * sizeof(double), hipMemcpyHostToDevice, F.st));
    C3D_HIP(hipMemcpyAsync(F.frgb.p, LF.rgb.data(), 3 * size_t(nf), hipMemcpyHostToDevice, F.st));
  }
  F.h_nline = nf;
  if (upload_table(F, F.hdr, offsetof(FrameHdr, nline), &F.h_nline, sizeof(uint32_t))) return 1;
  (void)fp64;
  if (F.frag_cap == 0) F.frag_cap = kDefaultFragCap;
  C3D_HIP(F.frags.ensure(size_t(F.frag_cap) * (fp64 ? sizeof(Frag64) : sizeof(Frag))));
  return 0;
}

// line fragments the buffers hold (the line kernels' grid in a replayed frame)
uint32_t line_capacity(const Frame& F) {
  const size_t n = std::min(std::min(F.fpix.bytes / sizeof(uint32_t), F.fdep.bytes / sizeof(double)), F.frgb.bytes / 3);
  return uint32_t(std::min<size_t>(n, 0x7fffffff));
}

// z-buffer passes, lines and resolve of every image over the mesh M;
// `lines` sizes the line kernels' grid (at least the frame's line fragments)
template <typename R>
int raster_launch(Lib&, Frame& F, const contour3d_amd_mesh& M, const RasterJob<R>* rjobs, int njobs, uint64_t npx,
                  uint32_t rgrid, uint32_t lines) {
  const bool fp64 = sizeof(R) == 8;
  FrameHdr* hdr = F.hdr.as<FrameHdr>();
  uint32_t* id = fp64 ? F.id.as<uint32_t>() : nullptr;
  uint64_t* key = F.key.as<uint64_t>();
  const MeshView<R> mv = view_of<R>(M);
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_RASTER)) return 1;
  C3D_HIP(launch_zinit(key, id, false, npx, fp64 ? 0 : 1, false, hdr, F.st));
  if constexpr (sizeof(R) == 8) {
    // depth pass keeping the fragments that may end up nearest, then the id
    // pass over that list; when the list overflows, the geometry id pass
    // runs instead (both decide on the device) and the list grows for the
    // next frame
    C3D_HIP(launch_raster<R>(mv, rjobs, njobs, hdr, kRasterDepth64, key, id, F.frags.as<Frag>(), F.frag_cap, rgrid,
                             F.st));
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 0, 0, key, id, false, F.st));
    C3D_HIP(launch_frag_ids(F.frags.p, false, F.frag_cap, hdr, key, id,
                            std::min<uint32_t>((F.frag_cap + 255) / 256, 4096), F.st));
    C3D_HIP(launch_raster<R>(mv, rjobs, njobs, hdr, kRasterId64, key, id, nullptr, 0, uint32_t(persistent_blocks()),
                             F.st));
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 1, 0, key, id, false, F.st));
  } else {
    C3D_HIP(launch_raster<R>(mv, rjobs, njobs, hdr, kRasterFp32, key, id, nullptr, 0, rgrid, F.st));
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 0, 1, key, id, false, F.st));
  }
  if (prof_end(F)) return 1;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_RESOLVE)) return 1;
  C3D_HIP(launch_resolve<R>(mv, rjobs, njobs, npx, M.meta.as<ObjMeta>(), int(M.base.size()), F.frgb.as<uint8_t>(),
                            hdr, key, id, 1u << 30, F.st));
  if (prof_end(F)) return 1;
  return 0;
}

// raster workgroups: FP64 rasterises with persistent workgroups (64 per CU)
// that overlap the next round's loads with drawing, FP32 with one workgroup
// per round unless `persistent`
uint32_t raster_grid(const Lib& L, bool fp64, uint64_t nwork, bool persistent) {
  uint64_t rg = (nwork + kThreads - 1) / kThreads;
  const uint64_t rcap = L.raster_cap ? L.raster_cap : ((fp64 || persistent) ? uint64_t(persistent_blocks()) * 8 : 0);
  if (rcap && rg > rcap) rg = rcap;
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(rg, 1u << 30)));
}

// copies to host outputs and the header, then the frame's completion event
int raster_copies(Frame& F, const std::vector<Image>& imgs, bool fp64) {
  for (const Image& im : imgs) {
    const uint64_t n = uint64_t(im.hv.W) * uint64_t(im.hv.H);
    if (!im.rgb_dev)
      C3D_HIP(hipMemcpyAsync(im.rgb, F.rgb.as<uint8_t>() + 3 * im.px0, 3 * n, hipMemcpyDeviceToHost, F.st));
    if (im.depth && !im.dep_dev)
      C3D_HIP(hipMemcpyAsync(im.depth, F.depth.as<double>() + im.px0, n * sizeof(double), hipMemcpyDeviceToHost,
                             F.st));
  }
  C3D_HIP(hipMemcpyAsync(F.hstage.p, F.hdr.p, sizeof(FrameHdr), hipMemcpyDeviceToHost, F.st));
  C3D_HIP(hipEventRecord(F.done, F.st));
  F.fp64 = fp64;
  return 0;
}

// z-buffer passes, lines and resolve of every image of a frame over the
// mesh M, then the copies to host outputs; frame_finish waits for them.
template <typename R>
int raster_frame(Lib& L, Frame& F, const contour3d_amd_mesh& M, std::vector<Image>& imgs, const RasterJob<R>* rjobs,
                 uint64_t npx, const FrameHdr& hh) {
  const uint32_t rgrid = raster_grid(L, sizeof(R) == 8, hh.nwork, false);
  if (raster_launch<R>(L, F, M, rjobs, int(imgs.size()), npx, rgrid, F.h_nline)) return 1;
  return raster_copies(F, imgs, sizeof(R) == 8);
}

void release_geoms(Frame& F) {
  for (Geom* g : F.geoms_used) g->busy--;
  F.geoms_used.clear();
}

// Waits for a frame's device work and host copies; grows the fragment list
// after an overflow.
int frame_finish(Frame& F) {
  const hipError_t e = hipEventSynchronize(F.done);
  release_geoms(F);
  C3D_HIP(e);
  FrameHdr& fin = F.last;
  std::memcpy(&fin, F.hstage.p, sizeof(FrameHdr));
  if (F.unread && !(fin.overflow & 1u)) {
    // totals of a frame planned without a read back, from its final header
    F.mesh.total = fin.ntri;
    F.prof.ntri += fin.ntri;
  }
  F.unread = false;
  if ((fin.overflow & 2u) && fin.nfrag > F.frag_cap && !F.noalloc) {
    const uint64_t grow = uint64_t(fin.nfrag) + fin.nfrag / 4;
    F.frag_cap = uint32_t(grow > 0xffffffffull ? 0xffffffffull : grow);
  }
  return prof_collect(F);
}

template <typename R>
int build_mesh_t(Lib& L, Frame& F, contour3d_amd_mesh& M, const contour3d_contour_obj_t* objs, size_t n) {
  std::vector<const contour3d_contour_obj_t*> v(n);
  for (size_t o = 0; o < n; o++) v[o] = objs + o;
  FrameHdr hh{};
  return tessellate<R>(L, F, M, v, nullptr, 0, hh);
}

int build_mesh(Lib& L, Frame& F, contour3d_amd_mesh& M, const contour3d_contour_obj_t* objs, size_t n, int precision) {
  if (n && !objs) {
    log_error("contour objects pointer is NULL");
    return 1;
  }
  if (precision == CONTOUR3D_AMD_FP64) return build_mesh_t<double>(L, F, M, objs, n);
  if (precision == CONTOUR3D_AMD_FP32) return build_mesh_t<float>(L, F, M, objs, n);
  log_error("unknown precision %d", precision);
  return 1;
}

// Validates the images, places them in the frame's pixel space and sets up
// their output targets.
int frame_setup(Frame& F, std::vector<Image>& imgs, uint64_t& npx) {
  npx = 0;
  for (size_t q = 0; q < imgs.size(); q++) {
    Image& im = imgs[q];
    const HostView& hv = im.hv;
    if (hv.W == 0 || hv.H == 0 || hv.W > (1u << 16) || hv.H > (1u << 16) ||
        uint64_t(hv.W) * uint64_t(hv.H) >= (uint64_t(1) << 31)) {
      log_error("screen size %zu x %zu out of range", hv.W, hv.H);
      return 1;
    }
    if (!im.rgb) {
      log_error("rgb output pointer is NULL");
      return 1;
    }
    if (im.nobj && !im.objs) {
      log_error("contour objects pointer is NULL");
      return 1;
    }
    // every image starts on a workgroup boundary of the frame's pixels
    im.px0 = (npx + kThreads - 1) / kThreads * kThreads;
    npx = im.px0 + uint64_t(hv.W) * uint64_t(hv.H);
  }
  if (npx >= (uint64_t(1) << 31)) {
    log_error("%llu pixels in one frame exceed 2^31", (unsigned long long)npx);
    return 1;
  }
  bool any_host_rgb = false;
  for (const Image& im : imgs) any_host_rgb = any_host_rgb || !im.rgb_dev;
  if (any_host_rgb) C3D_HIP(F.rgb.ensure(3 * npx + 1));

  bool any_dep = false;
  for (const Image& im : imgs) any_dep = any_dep || (im.depth && !im.dep_dev);
  if (any_dep) C3D_HIP(F.depth.ensure(npx * sizeof(double)));
  return 0;
}

template <typename R>
void fill_rjobs(Frame& F, const std::vector<Image>& imgs) {
  std::vector<RasterJob<R>>& rj = F.rjobs_host<R>();
  rj.assign(imgs.size(), RasterJob<R>{});
  for (size_t q = 0; q < imgs.size(); q++) {
    const Image& im = imgs[q];
    RasterJob<R>& j = rj[q];
    j.v = dev_view<R>(im.hv);
    j.px0 = im.px0;
    j.obj0 = im.obj0;
    j.obj1 = im.obj1;
    j.rgb = im.rgb_dev ? im.rgb : F.rgb.as<uint8_t>() + 3 * im.px0;
    j.depth = im.depth ? (im.dep_dev ? im.depth : F.depth.as<double>() + im.px0) : nullptr;
    j.bg[0] = im.hv.bg.r;
    j.bg[1] = im.hv.bg.g;
    j.bg[2] = im.hv.bg.b;
  }
}

// Uploads, for every field pass of the frame, the images that show any of
// its objects: offsets [njobs + 1] into the same array, then image indices.
int upload_pass_images(Frame& F, const std::vector<Image>& imgs) {
  const size_t nj = F.h_tjobs.size();
  std::vector<uint32_t>& J = F.h_jimg;
  J.assign(nj + 1, 0);
  for (size_t j = 0; j < nj; j++) {
    J[j] = uint32_t(J.size());
    const ThrSet& ts = F.h_tjobs[j].ts;
    for (size_t k = 0; k < imgs.size(); k++) {
      bool shows = false;
      for (int q = 0; q < ts.n && !shows; q++) shows = ts.obj[q] >= imgs[k].obj0 && ts.obj[q] < imgs[k].obj1;
      if (shows) J.push_back(uint32_t(k));
    }
  }
  J[nj] = uint32_t(J.size());
  C3D_HIP(F.jimg.ensure(J.size() * sizeof(uint32_t)));
  return upload_table(F, F.jimg, 0, J.data(), J.size() * sizeof(uint32_t));
}

// A frame without a stored mesh: z-buffer clear, the fused tessellation and
// raster, lines and the FP64 id passes, and the resolve from the fields.
// `lines` sizes the line kernels' grid (at least the frame's fragments).
template <typename R>
int fused_launch(Lib& L, Frame& F, const contour3d_amd_mesh& M, const RasterJob<R>* rjobs, int njobs, uint64_t npx,
                 uint32_t lines) {
  const bool fp64 = sizeof(R) == 8;
  FrameHdr* hdr = F.hdr.as<FrameHdr>();
  void* id = fp64 ? F.id.p : nullptr;
  uint64_t* key = F.key.as<uint64_t>();
  const TessLayout& tl = F.tl;
  const TessJob* tj = F.tjobs.as<TessJob>();
  const uint32_t* jimg = F.jimg.as<uint32_t>();
  // FP64 frames name triangles by object, cell and slot: no tile starts
  // and no cell store
  const uint64_t* start = fp64 ? nullptr : M.bstart.as<uint64_t>();
  uint32_t* cell = fp64 ? nullptr : M.cell.as<uint32_t>();
  const uint32_t grid = L.tess_grid(tl.ntj);
  const uint32_t fgrid = std::min<uint32_t>((F.frag_cap + 255) / 256, 4096);
  void* frags = F.frags.p;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_RASTER)) return 1;
  C3D_HIP(launch_zinit(key, id, fp64, npx, fp64 ? 0 : 1, fp64, hdr, F.st));
  if constexpr (sizeof(R) == 8) {
    // every covered pixel into the fragment list, then its keys; after an
    // overflow of the list (which grows for the next frame) the z-test
    // passes over the geometry run instead (each kernel decides on the
    // device)
    if (prof_inner_begin(F, CONTOUR3D_AMD_STAGE_FUSED)) return 1;
    C3D_HIP(launch_tess_raster<R>(tj, tl.njobs, tl.ntj, start, cell, 0, rjobs, jimg, hdr, kRasterAppend, key, id, frags,
                                  F.frag_cap, grid, F.st));
    if (prof_inner_end(F)) return 1;
    C3D_HIP(launch_frag_keys(frags, true, F.frag_cap, hdr, key, fgrid, F.st));
    C3D_HIP(launch_tess_raster<R>(tj, tl.njobs, tl.ntj, start, cell, 0, rjobs, jimg, hdr, kRasterKey64, key, id,
                                  nullptr, 0, grid, F.st));
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 0, 0, key, id, true, F.st));
    C3D_HIP(launch_frag_ids(frags, true, F.frag_cap, hdr, key, id, fgrid, F.st));
    C3D_HIP(launch_tess_raster<R>(tj, tl.njobs, tl.ntj, start, cell, 0, rjobs, jimg, hdr, kRasterId64, key, id,
                                  nullptr, 0, grid, F.st));
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 1, 0, key, id, true, F.st));
  } else {
    // one pass: the packed key carries the id
    if (prof_inner_begin(F, CONTOUR3D_AMD_STAGE_FUSED)) return 1;
    C3D_HIP(launch_tess_raster<R>(tj, tl.njobs, tl.ntj, start, cell, M.cap, rjobs, jimg, hdr, kRasterFp32, key, id,
                                  nullptr, 0, grid, F.st));
    if (prof_inner_end(F)) return 1;
    C3D_HIP(launch_line_frags(F.fpix.as<uint32_t>(), F.fdep.as<double>(), lines, hdr, 0, 1, key, id, false, F.st));
  }
  if (prof_end(F)) return 1;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_RESOLVE)) return 1;
  C3D_HIP(launch_resolve_field<R>(tj, cell, rjobs, njobs, npx, M.meta.as<ObjMeta>(), int(M.base.size()),
                                  F.frgb.as<uint8_t>(), hdr, key, id, F.st));
  if (prof_end(F)) return 1;
  return 0;
}

// The frame's launch sequence (count ... resolve) as a graph, replayed
// while its layout stays the same: the kernels' arguments and grids depend
// only on the layout, and everything else (fields, thresholds, colours,
// cameras, line fragments, outputs) is read from the device tables uploaded
// before each launch.  The plan kernel flags a frame whose triangles exceed
// the capacity reserved by earlier frames; the caller then renders it again
// eagerly.  The first frame of a layout, and one whose line fragments
// outgrow their buffers, runs eagerly.
template <typename R>
int graph_frame(Lib& L, Frame& F, contour3d_amd_mesh& M, std::vector<Image>& imgs,
                const std::vector<const contour3d_contour_obj_t*>& objs, RasterJob<R>* rjd, uint64_t npx,
                bool& replayed) {
  const bool fp64 = sizeof(R) == 8;
  const int njobs = int(imgs.size());
  if (tess_prepare<R>(L, F, M, objs)) return 1;
  if (upload_pass_images(F, imgs)) return 1;
  if (raster_prepare<R>(F, npx)) return 1;
  // FP64 frames have no cell store to outgrow
  const bool ready =
      (fp64 || (M.cap > 0 && M.cell.p && M.cap < kMaxFrameTriangles)) && F.h_nline <= line_capacity(F);
  if (!ready) {
    F.drop_graph();
    if (!fp64) {
      FrameHdr hh{};
      if (tess_head<R>(L, F, M, rjd, njobs, ~uint64_t(0))) return 1;
      if (tess_readback<R>(F, M, hh, F.h_nline, true)) return 1;
    } else {
      F.unread = true;
    }
    if (fused_launch<R>(L, F, M, rjd, njobs, npx, F.h_nline)) return 1;
    return raster_copies(F, imgs, fp64);
  }
  const uint32_t lines = line_capacity(F);
  const std::vector<uint64_t> sig = {
      sizeof(R), uint64_t(njobs), npx, F.tl.nbt, F.tl.ntj, uint64_t(F.tl.njobs), uint64_t(F.tl.nobj),
      L.tess_cap, L.tess_grid(F.tl.ntj), L.tess_rows, L.tess_zc, lines, M.cap, F.frag_cap,
      uint64_t(uintptr_t(F.cnt.p)), uint64_t(uintptr_t(F.tiles.p)), uint64_t(uintptr_t(F.tjobs.p)),
      uint64_t(uintptr_t(F.rjobs.p)), uint64_t(uintptr_t(F.hdr.p)), uint64_t(uintptr_t(F.jimg.p)),
      uint64_t(uintptr_t(F.key.p)), uint64_t(uintptr_t(F.id.p)), uint64_t(uintptr_t(F.frags.p)),
      uint64_t(uintptr_t(F.fpix.p)), uint64_t(uintptr_t(F.fdep.p)), uint64_t(uintptr_t(F.frgb.p)),
      uint64_t(uintptr_t(M.bstart.p)), uint64_t(uintptr_t(M.meta.p)), uint64_t(uintptr_t(M.cell.p))};
  if (!F.gexec || F.gsig != sig) {
    F.drop_graph();
    (void)persistent_blocks();  // device query outside the capture
    const bool prof_on = F.prof.on;
    F.prof.on = false;
    C3D_HIP(hipStreamBeginCapture(F.st, hipStreamCaptureModeThreadLocal));
    int rc = fp64 ? 0 : tess_head<R>(L, F, M, rjd, njobs, M.cap);
    if (!rc) rc = fused_launch<R>(L, F, M, rjd, njobs, npx, lines);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(F.st, &g);
    F.prof.on = prof_on;
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      log_error("graph capture failed: %s", hipGetErrorString(e));
      return 1;
    }
    const hipError_t ei = hipGraphInstantiate(&F.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      F.gexec = nullptr;
      log_error("graph instantiation failed: %s", hipGetErrorString(ei));
      return 1;
    }
    F.gsig = sig;
  }
  C3D_HIP(hipGraphLaunch(F.gexec, F.st));
  replayed = true;
  F.unread = true;
  return raster_copies(F, imgs, fp64);
}

// Tessellate the objects of every image and render them all.
template <typename R>
int run_frame_t(Lib& L, Frame& F, contour3d_amd_mesh& M, std::vector<Image>& imgs, bool graph, bool& replayed,
                bool speculate) {
  replayed = false;
  F.unread = false;
  uint64_t npx = 0;
  if (frame_setup(F, imgs, npx)) return 1;
  // the frame's object list: each image's objects, unless an earlier image
  // shows exactly the same ones
  std::vector<const contour3d_contour_obj_t*> objs;
  for (size_t q = 0; q < imgs.size(); q++) {
    Image& im = imgs[q];
    bool shared = false;
    for (size_t p = 0; p < q && !shared; p++) {
      const Image& ip = imgs[p];
      if (ip.nobj != im.nobj) continue;
      bool eq = true;
      for (size_t o = 0; o < im.nobj && eq; o++) eq = same_obj(ip.objs[o], im.objs[o]);
      if (eq) {
        im.obj0 = ip.obj0;
        im.obj1 = ip.obj1;
        shared = true;
      }
    }
    if (shared) continue;
    im.obj0 = uint32_t(objs.size());
    for (size_t o = 0; o < im.nobj; o++) objs.push_back(im.objs + o);
    im.obj1 = uint32_t(objs.size());
  }
  if (frame_lines(F, imgs)) return 1;
  fill_rjobs<R>(F, imgs);
  C3D_HIP(F.rjobs.ensure(imgs.size() * sizeof(RasterJob<R>)));
  if (upload_table(F, F.rjobs, 0, F.rjobs_host<R>().data(), imgs.size() * sizeof(RasterJob<R>))) return 1;
  RasterJob<R>* rjd = F.rjobs.as<RasterJob<R>>();
  const int njobs = int(imgs.size());
  // FP64 ids hold the object index below kLineObj - 1 (tri_id64)
  if (sizeof(R) == 8 && objs.size() + 1 >= kLineObj) {
    log_error("%zu contour objects exceed a frame's %llu", objs.size(), (unsigned long long)(kLineObj - 2));
    return 1;
  }
  if (graph) return graph_frame<R>(L, F, M, imgs, objs, rjd, npx, replayed);
  FrameHdr hh{};
  if (tess_prepare<R>(L, F, M, objs)) return 1;
  if (upload_pass_images(F, imgs)) return 1;
  // the line fragments' count goes to the header before the plan reads it
  if (raster_prepare<R>(F, npx)) return 1;
  if constexpr (sizeof(R) == 8) {
    // no count, scan or plan: triangles are named by object, cell and slot
    // (tri_id64), and the frame's triangle total comes with its final
    // header
    if (fused_launch<R>(L, F, M, rjd, njobs, npx, F.h_nline)) return 1;
    replayed = true;
    F.unread = true;
    return raster_copies(F, imgs, true);
  }
  if (speculate && M.cap > 0 && M.cell.p && M.cap < kMaxFrameTriangles) {
    // planned against the cell store that earlier frames reserved, with no
    // read back of the totals (no synchronisation mid-frame); the plan
    // flags a frame that needs more, and run_frame renders it again
    if (tess_head<R>(L, F, M, rjd, njobs, M.cap)) return 1;
    if (fused_launch<R>(L, F, M, rjd, njobs, npx, F.h_nline)) return 1;
    replayed = true;
    F.unread = true;
    return raster_copies(F, imgs, sizeof(R) == 8);
  }
  if (tess_head<R>(L, F, M, rjd, njobs, ~uint64_t(0))) return 1;
  if (tess_readback<R>(F, M, hh, F.h_nline, true)) return 1;
  if (fused_launch<R>(L, F, M, rjd, njobs, npx, F.h_nline)) return 1;
  return raster_copies(F, imgs, sizeof(R) == 8);
}

// Renders the images as one frame into F; with wait, also waits for the
// frame (otherwise frame_finish must be called before F is reused).  With
// graphs on, a waited frame replays the captured launch sequence of its
// layout; a replay whose triangles outgrew the reserved capacity is
// rendered again eagerly.
int run_frame(Lib& L, Frame& F, std::vector<Image>& imgs, int precision, bool wait = true) {
  if (precision != CONTOUR3D_AMD_FP64 && precision != CONTOUR3D_AMD_FP32) {
    log_error("unknown precision %d", precision);
    return 1;
  }
  const bool graph = wait && L.graphs;
  for (int attempt = 0; attempt < 2; attempt++) {
    bool replayed = false;
    const bool g = graph && attempt == 0;
    // a waited frame skips the mid-frame read back while the reserved
    // capacity holds it; a frame that outgrows it is rendered again
    const bool spec = wait && attempt == 0;
    const int rc = precision == CONTOUR3D_AMD_FP64 ? run_frame_t<double>(L, F, F.mesh, imgs, g, replayed, spec)
                                                   : run_frame_t<float>(L, F, F.mesh, imgs, g, replayed, spec);
    if (rc) {
      (void)hipStreamSynchronize(F.st);
      release_geoms(F);
      F.drop_graph();
      return rc;
    }
    if (!wait) return 0;
    if (frame_finish(F)) {
      F.drop_graph();
      return 1;
    }
    if (!(replayed && (F.last.overflow & 1u))) return 0;
    F.drop_graph();  // the cell store must grow: render again, with the read back
  }
  return 0;
}

// The allocation-free frame: planned against the scratch array's cell
// store with no read back of its totals, never re-rendered or grown; a
// frame beyond a scratch maximum fails with an error.
int run_frame_noalloc(Lib& L, Frame& F, std::vector<Image>& imgs, int precision) {
  if (precision != CONTOUR3D_AMD_FP64 && precision != CONTOUR3D_AMD_FP32) {
    log_error("unknown precision %d", precision);
    return 1;
  }
  bool replayed = false;
  const int rc = precision == CONTOUR3D_AMD_FP64 ? run_frame_t<double>(L, F, F.mesh, imgs, false, replayed, true)
                                                 : run_frame_t<float>(L, F, F.mesh, imgs, false, replayed, true);
  if (rc) {
    (void)hipStreamSynchronize(F.st);
    return rc;
  }
  if (!replayed) {
    log_error("allocation-free frame not planned against the scratch array");
    return 1;
  }
  if (frame_finish(F)) return 1;
  if (F.last.overflow & 1u) {
    log_error("the allocation-free call holds at most %llu triangles (and 2^32 triangle and line ids); the frame "
              "has %llu triangles", (unsigned long long)F.mesh.cap, (unsigned long long)F.last.ntri);
    return 1;
  }
  return 0;
}

// Render an already built mesh (all its objects) as one image.
template <typename R>
int render_mesh_t(Lib& L, Frame& F, const contour3d_amd_mesh& M, std::vector<Image>& imgs) {
  uint64_t npx = 0;
  if (frame_setup(F, imgs, npx)) return 1;
  for (Image& im : imgs) {
    im.obj0 = 0;
    im.obj1 = uint32_t(M.base.size());
  }
  if (frame_lines(F, imgs)) return 1;
  fill_rjobs<R>(F, imgs);
  C3D_HIP(F.rjobs.ensure(imgs.size() * sizeof(RasterJob<R>)));
  if (upload_table(F, F.rjobs, 0, F.rjobs_host<R>().data(), imgs.size() * sizeof(RasterJob<R>))) return 1;
  RasterJob<R>* rjd = F.rjobs.as<RasterJob<R>>();
  C3D_HIP(F.hdr.ensure(sizeof(FrameHdr)));
  C3D_HIP(F.hstage.ensure(sizeof(FrameHdr)));
  if (!frame_ids_fit(M.total, F.h_frags.pix.size())) {
    log_error("%llu triangles and %zu line fragments exceed the 2^32 ids of one frame",
              (unsigned long long)M.total, F.h_frags.pix.size());
    return 1;
  }
  // the line fragments' count goes to the header before the plan reads it
  if (raster_prepare<R>(F, npx)) return 1;
  // the plan kernel recomputes the ranges from the mesh's tile starts
  C3D_HIP(launch_plan<R>(M.bstart.as<uint64_t>(), M.ntiles, M.meta.as<ObjMeta>(), int(M.base.size()), rjd,
                         int(imgs.size()), F.hdr.as<FrameHdr>(), ~uint64_t(0), F.st));
  FrameHdr hh{};
  hh.ntri = M.total;
  hh.nwork = (M.total + kThreads) * imgs.size();  // bound on the aligned work space
  return raster_frame<R>(L, F, M, imgs, rjd, npx, hh);
}

int render_mesh(Lib& L, Frame& F, const contour3d_amd_mesh& M, std::vector<Image>& imgs) {
  const int rc = M.precision == CONTOUR3D_AMD_FP64 ? render_mesh_t<double>(L, F, M, imgs)
                                                   : render_mesh_t<float>(L, F, M, imgs);
  if (rc) {
    (void)hipStreamSynchronize(F.st);
    return rc;
  }
  return frame_finish(F);
}

Image make_image(const V* cam, const V* light, const size_t sizes[2], const V* ctr, const V local[2],
                 const contour3d_color_t* bg, size_t nobj, const contour3d_contour_obj_t* objs, size_t nlines,
                 const contour3d_line_obj_t* lines, uint8_t* rgb, double* depth, bool rgb_host_only) {
  Image im;
  im.hv = make_view(cam, light, sizes, ctr, local, bg);
  im.nobj = nobj;
  im.objs = objs;
  im.nlines = nlines;
  im.lines = lines;
  im.rgb = rgb;
  im.depth = depth;
  im.rgb_dev = !rgb_host_only && rgb && is_device_ptr(rgb);
  im.dep_dev = depth && is_device_ptr(depth);
  return im;
}

// ----------------------------------------------------------------------
// serving loop
//
// Requests (one image each) queue in arrival order.  A worker thread forms
// a batch from the head of the queue: requests of the head's arithmetic
// form, up to max_images, while the estimated device bytes of the frames
// in flight stay within the HBM budget.  Two frames alternate, each with
// its own stream and buffers: while one frame runs on the device, the next
// batch is prepared and enqueued on the other, so requests that arrive
// during a frame start as soon as a frame slot and budget free up.  A host
// function at the end of each frame's stream wakes the worker.

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Request {
  uint64_t ticket = 0;
  contour3d_amd_image_t img{};
  int precision = 0;
  uint64_t bytes = 0;  // admission estimate
  double t_submit = 0, t_done = 0;
  int status = -1;  // -1 pending, 0 done, 1 failed
  std::string fname;          // PPM output, written by the server's writer thread
  uint8_t* pinned = nullptr;  // the pinned host image it is rendered into
  size_t pinned_bytes = 0;
  // snapshot_fields: the request's own contour objects, pointing at device
  // copies of its fields taken at submit (returned to the pool at completion)
  std::vector<contour3d_contour_obj_t> objs;
  std::vector<std::pair<void*, size_t>> snaps;
};

struct Flight {
  Server* srv = nullptr;
  int slot = 0;
  std::vector<std::shared_ptr<Request>> reqs;
  uint64_t bytes = 0;
  double cells = 0;
  double triangles = 0;  // of the frame, known once it is enqueued
  bool done = false;     // set by the stream's host function
};

struct Server {
  contour3d_amd_server_opts_t opts{};
  std::thread worker;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
  std::deque<std::unique_ptr<Flight>> flights;
  uint64_t next_ticket = 1;
  bool stop = false;
  Frame frames[2];
  double tri_per_cell = 0.5;  // running estimate of triangles per cell and threshold
  uint64_t admitted = 0;      // estimated bytes of the frames in flight
  std::vector<double> lat_ms;
  uint64_t completed = 0, failed = 0, batches = 0, images = 0, peak = 0;
  double t_first = 0, t_last = 0;
  // PPM requests: rendered frames wait here for the writer thread
  std::thread writer;
  std::condition_variable cv_write;
  std::deque<std::shared_ptr<Request>> to_write;
  bool writer_stop = false;
  // pinned host images of PPM requests, reused (allocated with hipHostMalloc
  // by submit, freed by stop)
  std::mutex pool_mu;
  std::vector<std::pair<uint8_t*, size_t>> pinned_free, pinned_all;
  // device snapshots of fields (snapshot_fields), reused: allocated by
  // submit, taken and returned under pool_mu, freed by stop; snap_held counts
  // the bytes that requests hold, within the HBM budget
  std::vector<std::pair<void*, size_t>> snap_free, snap_all;
  uint64_t snap_held = 0;
  hipStream_t copy_st = nullptr;  // the snapshot copies
};

uint8_t* pinned_take(Server& S, size_t bytes, size_t& got) {
  {
    std::lock_guard<std::mutex> g(S.pool_mu);
    size_t best = S.pinned_free.size();
    for (size_t k = 0; k < S.pinned_free.size(); k++)
      if (S.pinned_free[k].second >= bytes && (best == S.pinned_free.size() || S.pinned_free[k].second <
                                                                                    S.pinned_free[best].second))
        best = k;
    if (best < S.pinned_free.size()) {
      const auto b = S.pinned_free[best];
      S.pinned_free.erase(S.pinned_free.begin() + long(best));
      got = b.second;
      return b.first;
    }
  }
  void* p = nullptr;
  if (injected_alloc_failure() || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) return nullptr;
  std::lock_guard<std::mutex> g(S.pool_mu);
  S.pinned_all.push_back({static_cast<uint8_t*>(p), bytes});
  got = bytes;
  return static_cast<uint8_t*>(p);
}

void pinned_give(Server& S, uint8_t* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> g(S.pool_mu);
  S.pinned_free.push_back({p, bytes});
}

// A device buffer of at least `bytes` for a field snapshot: the smallest
// free one that fits, else a new one; null when the snapshots held would
// exceed the HBM budget or the allocation fails.
void* snap_take(Server& S, size_t bytes, size_t& got) {
  {
    std::lock_guard<std::mutex> g(S.pool_mu);
    if (S.snap_held + bytes > S.opts.hbm_budget) return nullptr;
    size_t best = S.snap_free.size();
    for (size_t k = 0; k < S.snap_free.size(); k++)
      if (S.snap_free[k].second >= bytes &&
          (best == S.snap_free.size() || S.snap_free[k].second < S.snap_free[best].second))
        best = k;
    if (best < S.snap_free.size()) {
      const auto b = S.snap_free[best];
      S.snap_free.erase(S.snap_free.begin() + long(best));
      S.snap_held += b.second;
      got = b.second;
      return b.first;
    }
  }
  void* p = nullptr;
  if (injected_alloc_failure() || hipMalloc(&p, bytes) != hipSuccess || !p) return nullptr;
  std::lock_guard<std::mutex> g(S.pool_mu);
  S.snap_all.push_back({p, bytes});
  S.snap_held += bytes;
  got = bytes;
  return p;
}

// Returns a request's snapshots to the pool (no HIP call: runs in the
// stream's host function too).
void snap_give(Server& S, Request& r) {
  if (r.snaps.empty()) return;
  std::lock_guard<std::mutex> g(S.pool_mu);
  for (auto& b : r.snaps) {
    S.snap_free.push_back(b);
    S.snap_held -= b.second;
  }
  r.snaps.clear();
}

std::mutex g_server_mu;  // server start/stop against submit/wait/stats

double obj_cells(const contour3d_contour_obj_t& c) {
  double n = 1;
  for (int d = 0; d < 3; d++) n *= c.glsizes[d] > 1 ? double(c.glsizes[d] - 1) : 0.0;
  return n;
}

// device bytes a request adds to a frame: the cell index of each triangle
// (a frame stores no triangles), z-buffer and image buffers, an FP64
// fragment-list share (four covered pixels per pixel), and host fields to
// upload
uint64_t estimate_bytes(const Server& S, const contour3d_amd_image_t& im, int precision) {
  double tri = 0, upload = 0;
  for (size_t o = 0; o < im.num_contours; o++) {
    const contour3d_contour_obj_t& c = im.contour_objs[o];
    tri += obj_cells(c) * S.tri_per_cell;
    if (!is_device_ptr(c.array)) upload += 8.0 * double(c.glsizes[0]) * double(c.glsizes[1]) * double(c.glsizes[2]);
  }
  const double npx = double(im.screen_sizes[0]) * double(im.screen_sizes[1]);
  const double per_px = 8 + (precision == CONTOUR3D_AMD_FP64 ? 4 + 4 * 16 : 0) + 3 + (im.depth ? 8 : 0);
  return uint64_t(1.25 * tri * sizeof(uint32_t) + upload + npx * per_px) + (uint64_t(1) << 20);
}

void complete_locked(Server& S, Flight& f, int rc, double tri_total);

// Runs on a runtime thread once the frame's stream has passed its output
// copies: the requests' images are in their buffers, so they complete here
// (waiters wake at once); the worker then retires the frame.  Calls no HIP
// function.
void on_frame_done(void* p) {
  Flight* f = static_cast<Flight*>(p);
  std::lock_guard<std::mutex> g(f->srv->m);
  f->done = true;
  complete_locked(*f->srv, *f, 0, f->triangles);
  f->srv->cv_work.notify_all();
}

void complete(Server& S, Flight& f, int rc, double tri_total) {
  std::lock_guard<std::mutex> g(S.m);
  complete_locked(S, f, rc, tri_total);
}

void complete_locked(Server& S, Flight& f, int rc, double tri_total) {
  const double t = now_s();
  for (auto& r : f.reqs) {
    snap_give(S, *r);  // the frame that read the snapshots is done
    if (!rc && !r->fname.empty()) {
      // rendered into its pinned image: completes once the file is written
      S.to_write.push_back(r);
      S.cv_write.notify_one();
      continue;
    }
    if (r->pinned) pinned_give(S, r->pinned, r->pinned_bytes);
    r->pinned = nullptr;
    r->status = rc ? 1 : 0;
    r->t_done = t;
    if (rc) {
      S.failed++;
    } else {
      S.completed++;
      S.lat_ms.push_back((t - r->t_submit) * 1e3);
    }
  }
  if (!rc && f.cells > 0) {
    const double seen = tri_total / f.cells;
    S.tri_per_cell = std::max(seen, 0.8 * S.tri_per_cell + 0.2 * seen);
  }
  S.admitted -= f.bytes;
  S.batches++;
  S.images += f.reqs.size();
  S.t_last = t;
  S.cv_done.notify_all();
}

void server_loop(Lib& L, Server& S) {
  (void)hipSetDevice(L.device);
  std::unique_lock<std::mutex> lk(S.m);
  for (;;) {
    const double wait_s = S.opts.max_wait_ms * 1e-3;
    auto head_fits = [&] {
      return !S.queue.empty() && (S.flights.empty() || S.admitted + S.queue.front()->bytes <= S.opts.hbm_budget);
    };
    auto can_start = [&] {
      if (S.flights.size() >= 2 || !head_fits()) return false;
      // a lone request may wait a little for company while the device is idle
      return !S.flights.empty() || S.queue.size() >= S.opts.max_images || wait_s <= 0 ||
             now_s() >= S.queue.front()->t_submit + wait_s;
    };
    auto pred = [&] {
      return (!S.flights.empty() && S.flights.front()->done) || can_start() ||
             (S.stop && S.queue.empty() && S.flights.empty());
    };
    // while the device is idle, the oldest request's window bounds the wait
    wait_work(lk, S.cv_work, pred, [&](std::chrono::steady_clock::time_point& until) {
      if (S.queue.empty() || wait_s <= 0 || !S.flights.empty()) return false;
      until = std::chrono::steady_clock::now() +
              std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                  std::chrono::duration<double>(std::max(0.0, S.queue.front()->t_submit + wait_s - now_s())));
      return true;
    });
    if (!S.flights.empty() && S.flights.front()->done) {
      std::unique_ptr<Flight> f = std::move(S.flights.front());
      S.flights.pop_front();
      lk.unlock();
      // its requests completed in the stream's host function
      if (frame_finish(S.frames[f->slot])) log_error("server: a frame failed after its images were copied");
      lk.lock();
      continue;
    }
    if (can_start()) {
      auto f = std::make_unique<Flight>();
      f->srv = &S;
      const int prec = S.queue.front()->precision;
      uint64_t bytes = 0;
      for (auto it = S.queue.begin(); it != S.queue.end() && f->reqs.size() < S.opts.max_images;) {
        Request& r = **it;
        if (r.precision != prec) {
          ++it;
          continue;
        }
        // first come, first served: stop at the first request that does not fit
        if (!(f->reqs.empty() && S.flights.empty()) && S.admitted + bytes + r.bytes > S.opts.hbm_budget) break;
        bytes += r.bytes;
        f->reqs.push_back(*it);
        it = S.queue.erase(it);
      }
      f->bytes = bytes;
      S.admitted += bytes;
      S.peak = std::max(S.peak, S.admitted);
      // the frame slot that no flight holds (at most one is in flight here)
      f->slot = S.flights.empty() ? 0 : (S.flights.front()->slot ^ 1);
      Flight* fp = f.get();
      S.flights.push_back(std::move(f));
      lk.unlock();
      Frame& F = S.frames[fp->slot];
      std::vector<Image> imgs;
      imgs.reserve(fp->reqs.size());
      for (auto& r : fp->reqs) {
        const contour3d_amd_image_t& im = r->img;
        for (size_t o = 0; o < im.num_contours; o++) fp->cells += obj_cells(im.contour_objs[o]);
        imgs.push_back(make_image(&im.camera_position, &im.light_direction, im.screen_sizes, &im.screen_center,
                                  im.screen_local, &im.bg_color, im.num_contours, im.contour_objs, im.num_lines,
                                  im.line_objs, im.rgb, im.depth, false));
      }
      int rc;
      {
        std::lock_guard<std::mutex> g(L.mu);
        rc = run_frame(L, F, imgs, prec, false);
        fp->triangles = double(F.mesh.total);
      }
      if (!rc && hipLaunchHostFunc(F.st, on_frame_done, fp) != hipSuccess) {
        log_error("server: cannot schedule the completion of a frame");
        rc = 1;
        (void)frame_finish(F);
      }
      if (rc) {
        lk.lock();
        std::unique_ptr<Flight> f2;
        for (auto it = S.flights.begin(); it != S.flights.end(); ++it)
          if (it->get() == fp) {
            f2 = std::move(*it);
            S.flights.erase(it);
            break;
          }
        lk.unlock();
        if (f2) complete(S, *f2, 1, 0);
      }
      lk.lock();
      continue;
    }
    if (S.stop && S.queue.empty() && S.flights.empty()) break;
  }
}

int write_ppm(const char* fname, size_t W, size_t H, const uint8_t* rgb);

// Writes the PPM files of rendered requests, in completion order, while the
// worker renders the next frames; each request completes when its file is
// written.  Calls no HIP function.
void writer_loop(Server& S) {
  std::unique_lock<std::mutex> lk(S.m);
  for (;;) {
    S.cv_write.wait(lk, [&] { return !S.to_write.empty() || S.writer_stop; });
    if (S.to_write.empty()) break;
    std::shared_ptr<Request> r = S.to_write.front();
    S.to_write.pop_front();
    lk.unlock();
    const int rc = write_ppm(r->fname.c_str(), r->img.screen_sizes[0], r->img.screen_sizes[1], r->pinned);
    pinned_give(S, r->pinned, r->pinned_bytes);
    lk.lock();
    r->pinned = nullptr;
    r->status = rc ? 1 : 0;
    r->t_done = now_s();
    if (rc) {
      S.failed++;
    } else {
      S.completed++;
      S.lat_ms.push_back((r->t_done - r->t_submit) * 1e3);
    }
    S.t_last = r->t_done;
    S.cv_done.notify_all();
  }
}

// output.c:125-140
int write_ppm(const char* fname, size_t W, size_t H, const uint8_t* rgb) {
  errno = 0;
  FILE* fp = std::fopen(fname, "w");
  if (!fp) {
    log_error("%s: %s", fname, std::strerror(errno));
    return 1;
  }
  errno = 0;
  std::fprintf(fp, "P6\n%zu %zu\n255\n", W, H);
  if (errno != 0) {
    log_error("%s: %s", fname, std::strerror(errno));
    std::fclose(fp);
    return 1;
  }
  const size_t n = std::fwrite(rgb, 3, W * H, fp);
  if (n != W * H) {
    log_error("fwrite failed (%zu expected, %zu returned)", W * H, n);
    std::fclose(fp);
    return 1;
  }
  std::fclose(fp);
  return 0;
}

// Unit vertex normals of every triangle of object `obj` of mesh M into the
// device buffer out (9 values of the mesh's precision per triangle), timed
// as the SMOOTH stage, and waited for (normal.c:46-165).
int smooth_object(Frame& F, const contour3d_amd_mesh& M, size_t obj, void* out) {
  const uint64_t n = M.count[obj];
  if (n == 0) return 0;
  const ObjMeta* md = M.meta.as<ObjMeta>() + obj;
  if (prof_begin(F, CONTOUR3D_AMD_STAGE_SMOOTH)) return 1;
  // the cell starts of every object, once per build
  hipError_t e = hipSuccess;
  if (!M.starts_ready) {
    C3D_HIP(M.cstart.ensure(M.code_bytes * sizeof(uint16_t)));
    for (size_t o = 0; o < M.count.size() && e == hipSuccess; o++)
      if (M.count[o]) e = launch_cell_starts(M.code.as<uint8_t>(), M.meta.as<ObjMeta>() + o, M.tiles[o],
                                             M.cstart.as<uint16_t>(), F.st);
    if (e != hipSuccess) {
      log_error("cell starts: %s", hipGetErrorString(e));
      return 1;
    }
    M.starts_ready = true;
  }
  e = M.precision == CONTOUR3D_AMD_FP64
          ? launch_vertex_normals<double>(view_of<double>(M), md, M.cstart.as<uint16_t>(), n,
                                          static_cast<double*>(out), F.st)
          : launch_vertex_normals<float>(view_of<float>(M), md, M.cstart.as<uint16_t>(), n,
                                         static_cast<float*>(out), F.st);
  if (e != hipSuccess) {
    log_error("vertex normals: %s", hipGetErrorString(e));
    return 1;
  }
  if (prof_end(F)) return 1;
  e = hipStreamSynchronize(F.st);
  if (e != hipSuccess) {
    log_error("vertex normals failed on the device: %s", hipGetErrorString(e));
    return 1;
  }
  return prof_collect(F);
}

// The reference runs one pencil per MPI rank and reads each rank's local
// extent from SimpleDecomp (extend.c:158-181); this library renders every
// contour object's array as the whole field.  Under an MPI launch of more
// than one rank the arrays are pencils, so contour3d_execute refuses the
// call instead of reading past them.  Ranks are taken from the launchers'
// environment (Open MPI, MPICH/Hydra and PMI, MVAPICH);
// CONTOUR3D_AMD_WHOLE_FIELD=1 declares that every rank passes whole fields.
int mpi_ranks_from_env() {
  const char* wf = std::getenv("CONTOUR3D_AMD_WHOLE_FIELD");
  if (wf && std::strcmp(wf, "1") == 0) return 1;
  const char* vars[] = {"OMPI_COMM_WORLD_SIZE", "PMI_SIZE", "PMIX_SIZE", "MV2_COMM_WORLD_SIZE"};
  int n = 1;
  for (const char* v : vars) {
    const char* s = std::getenv(v);
    if (!s || !*s) continue;
    char* end = nullptr;
    const long r = std::strtol(s, &end, 10);
    if (end && *end == '\0' && r > n) n = int(std::min<long>(r, 1 << 30));
  }
  return n;
}

bool args_ok(const void* cam, const void* light, const size_t* sizes, const void* ctr, const void* local,
             const void* bg) {
  if (!cam || !light || !sizes || !ctr || !local || !bg) {
    log_error("camera, light, screen and background arguments must not be NULL");
    return false;
  }
  return true;
}

}  // namespace
}  // namespace c3d

using c3d::lib;

extern "C" {

contour3d_vector_t contour3d_amd_identity(const contour3d_vector_t orthogonal) { return orthogonal; }

int contour3d_amd_init(int device) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  return c3d::ensure_init(L, device);
}

int contour3d_amd_finalize(void) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (!L.ready) return 0;
  if (L.server) {
    c3d::log_error("finalize: stop the server first");
    return 1;
  }
  (void)hipStreamSynchronize(L.main.st);
  for (auto& gm : L.geoms) {
    gm->grids.release();
    gm->table.release();
  }
  L.geoms.clear();
  L.main.release();
  L.scratch_frame.release();
  L.ready = false;
  return 0;
}

int contour3d_amd_tune(int param, int64_t value) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (value < 0 || value > int64_t(0xffffffffu)) {
    c3d::log_error("tune: value %lld out of range", (long long)value);
    return 1;
  }
  switch (param) {
    case CONTOUR3D_AMD_TUNE_TESS_GRID:
      L.tess_cap = uint32_t(value);
      return 0;
    case CONTOUR3D_AMD_TUNE_RASTER_GRID:
      L.raster_cap = uint32_t(value);
      return 0;
    case CONTOUR3D_AMD_TUNE_FRAG_CAPACITY:
      L.main.frag_cap = value > 0 ? uint32_t(std::min<int64_t>(value, 0xffffffffll)) : c3d::kDefaultFragCap;
      return 0;
    case CONTOUR3D_AMD_TUNE_GRAPHS:
      L.graphs = value != 0;
      if (!L.graphs) L.main.drop_graph();
      return 0;
    case CONTOUR3D_AMD_TUNE_FAIL_ALLOC:
      c3d::g_fail_alloc.store(value);
      return 0;
    case CONTOUR3D_AMD_TUNE_TESS_ROWS:
      L.tess_rows = uint32_t(value);
      return 0;
    case CONTOUR3D_AMD_TUNE_TESS_ZPLANES:
      L.tess_zc = uint32_t(value);
      return 0;
    default:
      c3d::log_error("tune: unknown parameter %d", param);
      return 1;
  }
}

int contour3d_amd_set_precision(int precision) {
  if (precision != CONTOUR3D_AMD_FP64 && precision != CONTOUR3D_AMD_FP32) {
    c3d::log_error("unknown precision %d", precision);
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  L.precision = precision;
  return 0;
}

int contour3d_amd_execute_rgb(const contour3d_vector_t* camera_position, const contour3d_vector_t* light_direction,
                              const size_t screen_sizes[2], const contour3d_vector_t* screen_center,
                              const contour3d_vector_t screen_local[2], const contour3d_color_t* bg_color,
                              size_t num_contours, const contour3d_contour_obj_t* contour_objs, size_t num_lines,
                              const contour3d_line_obj_t* line_objs, uint8_t* rgb, double* depth, int precision) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  if (!c3d::args_ok(camera_position, light_direction, screen_sizes, screen_center, screen_local, bg_color))
    return 1;
  if (!rgb) {
    c3d::log_error("rgb output pointer is NULL");
    return 1;
  }
  std::vector<c3d::Image> imgs{c3d::make_image(camera_position, light_direction, screen_sizes, screen_center,
                                               screen_local, bg_color, num_contours, contour_objs, num_lines,
                                               line_objs, rgb, depth, false)};
  return c3d::run_frame(L, L.main, imgs, precision);
}

int contour3d_amd_execute_rgb_noalloc(const contour3d_vector_t* camera_position,
                                      const contour3d_vector_t* light_direction, const size_t screen_sizes[2],
                                      const contour3d_vector_t* screen_center, const contour3d_vector_t screen_local[2],
                                      const contour3d_color_t* bg_color, size_t num_contours,
                                      const contour3d_contour_obj_t* contour_objs, size_t num_lines,
                                      const contour3d_line_obj_t* line_objs, uint8_t* rgb, double* depth,
                                      int precision) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (!L.ready || L.device < 0 || L.device >= c3d::kMaxDevices) {
    c3d::log_error("execute_rgb_noalloc: call contour3d_amd_init first (it sets up the scratch array)");
    return 1;
  }
  if (!c3d::g_scratch[L.device] || !L.scratch_frame.noalloc) {
    c3d::log_error("execute_rgb_noalloc: the scratch array could not be set up on device %d", L.device);
    return 1;
  }
  if (!c3d::args_ok(camera_position, light_direction, screen_sizes, screen_center, screen_local, bg_color))
    return 1;
  if (!rgb || !c3d::is_device_ptr(rgb) || (depth && !c3d::is_device_ptr(depth))) {
    c3d::log_error("execute_rgb_noalloc: rgb (and depth, if given) must be device pointers");
    return 1;
  }
  const uint64_t npx = uint64_t(screen_sizes[0]) * uint64_t(screen_sizes[1]);
  if (npx > c3d::kScratchPixels) {
    c3d::log_error("execute_rgb_noalloc: %zu x %zu pixels exceed the maximum of %llu", screen_sizes[0],
                   screen_sizes[1], (unsigned long long)c3d::kScratchPixels);
    return 1;
  }
  std::vector<c3d::Image> imgs{c3d::make_image(camera_position, light_direction, screen_sizes, screen_center,
                                               screen_local, bg_color, num_contours, contour_objs, num_lines,
                                               line_objs, rgb, depth, false)};
  return c3d::run_frame_noalloc(L, L.scratch_frame, imgs, precision);
}

int contour3d_amd_execute_batch(size_t num_images, const contour3d_amd_image_t* images, int precision) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  if (num_images == 0) return 0;
  if (!images) {
    c3d::log_error("images pointer is NULL");
    return 1;
  }
  std::vector<c3d::Image> imgs;
  imgs.reserve(num_images);
  for (size_t q = 0; q < num_images; q++) {
    const contour3d_amd_image_t& im = images[q];
    imgs.push_back(c3d::make_image(&im.camera_position, &im.light_direction, im.screen_sizes, &im.screen_center,
                                   im.screen_local, &im.bg_color, im.num_contours, im.contour_objs, im.num_lines,
                                   im.line_objs, im.rgb, im.depth, false));
  }
  return c3d::run_frame(L, L.main, imgs, precision);
}

int contour3d_execute(const sdecomp_info_t* const sdecomp_info, const contour3d_vector_t* const camera_position,
                      const contour3d_vector_t* const camera_look_at,
                      const contour3d_vector_t* const light_direction, const size_t screen_sizes[2],
                      const contour3d_vector_t* const screen_center, const contour3d_vector_t screen_local[2],
                      const contour3d_color_t* const bg_color, const size_t num_contours,
                      const contour3d_contour_obj_t* const contour3d_contour_objs, const size_t num_lines,
                      const contour3d_line_obj_t* const contour3d_line_objs, const char fname[]) {
  (void)sdecomp_info;    // one GPU holds the whole field
  (void)camera_look_at;  // unused by the reference too (struct.h:53-58)
  const int ranks = c3d::mpi_ranks_from_env();
  if (ranks > 1) {
    c3d::log_error("%d MPI ranks: each rank's array is a pencil of the field, and this library renders whole "
                   "fields on one GPU (set CONTOUR3D_AMD_WHOLE_FIELD=1 if every rank passes its whole field)",
                   ranks);
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  if (!c3d::args_ok(camera_position, light_direction, screen_sizes, screen_center, screen_local, bg_color))
    return 1;
  if (!fname) {
    c3d::log_error("output file name is NULL");
    return 1;
  }
  const size_t W = screen_sizes[0], H = screen_sizes[1];
  if (W == 0 || H == 0 || W > (1u << 16) || H > (1u << 16)) {
    c3d::log_error("screen size %zu x %zu out of range", W, H);
    return 1;
  }
  if (L.main.hrgb.ensure(W * H * 3) != hipSuccess) {
    c3d::log_error("failed to allocate image buffer");
    return 1;
  }
  uint8_t* host = static_cast<uint8_t*>(L.main.hrgb.p);
  std::vector<c3d::Image> imgs{c3d::make_image(camera_position, light_direction, screen_sizes, screen_center,
                                               screen_local, bg_color, num_contours, contour3d_contour_objs,
                                               num_lines, contour3d_line_objs, host, nullptr, true)};
  if (c3d::run_frame(L, L.main, imgs, L.precision)) {
    c3d::log_error("image output failed");
    return 1;
  }
  return c3d::write_ppm(fname, W, H, host);
}

int contour3d_amd_mesh_build(const contour3d_contour_obj_t* contour_objs, size_t num_contours, int precision,
                             contour3d_amd_mesh_t** mesh) {
  if (!mesh) {
    c3d::log_error("mesh output pointer is NULL");
    return 1;
  }
  *mesh = nullptr;
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  auto M = std::make_unique<contour3d_amd_mesh>();
  const int rc = c3d::build_mesh(L, L.main, *M, contour_objs, num_contours, precision);
  const hipError_t e = hipStreamSynchronize(L.main.st);
  c3d::release_geoms(L.main);
  if (rc || e != hipSuccess) {
    if (!rc) c3d::log_error("mesh build failed on the device");
    M->release();  // a failed build keeps no device memory
    return 1;
  }
  if (c3d::prof_collect(L.main)) return 1;
  *mesh = M.release();
  return 0;
}

int contour3d_amd_mesh_count(const contour3d_amd_mesh_t* mesh, size_t obj, size_t* ntri) {
  if (!mesh || !ntri || obj >= mesh->count.size()) {
    c3d::log_error("mesh_count: bad arguments");
    return 1;
  }
  *ntri = size_t(mesh->count[obj]);
  return 0;
}

int contour3d_amd_mesh_copy(const contour3d_amd_mesh_t* mesh, size_t obj, double* pos, double* wfn, double* vn,
                            uint32_t* cell, uint32_t* edges) {
  if (!mesh || obj >= mesh->count.size()) {
    c3d::log_error("mesh_copy: bad arguments");
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  const uint64_t n = mesh->count[obj], b = mesh->base[obj];
  if (n == 0) return 0;
  const bool fp64 = mesh->precision == CONTOUR3D_AMD_FP64;
  const size_t rs = fp64 ? 8 : 4;
  // record sizes and the byte offset of the edge word in its record (the
  // gather record in FP64, the position record in FP32)
  const size_t PS = fp64 ? sizeof(c3d::MeshPos<double>) : sizeof(c3d::MeshPos<float>);
  const size_t AS = fp64 ? sizeof(c3d::MeshAux<double>) : sizeof(c3d::MeshAux<float>);
  auto real_at = [&](const unsigned char* q) {
    if (rs == 8) {
      double x;
      std::memcpy(&x, q, 8);
      return x;
    }
    float y;
    std::memcpy(&y, q, 4);
    return double(y);
  };
  // the object's records, unpacked field by field
  std::vector<unsigned char> prec, arec;
  if (pos || edges) prec.resize(n * PS);
  if (wfn || cell || edges) arec.resize(n * AS);
  if ((!prec.empty() && hipMemcpy(prec.data(), static_cast<const unsigned char*>(mesh->pos.p) + b * PS, n * PS,
                                  hipMemcpyDeviceToHost) != hipSuccess) ||
      (!arec.empty() && hipMemcpy(arec.data(), static_cast<const unsigned char*>(mesh->aux.p) + b * AS, n * AS,
                                  hipMemcpyDeviceToHost) != hipSuccess)) {
    c3d::log_error("mesh_copy: device copy failed");
    return 1;
  }
  for (uint64_t t = 0; t < n; t++) {
    const unsigned char* rp = prec.empty() ? nullptr : prec.data() + t * PS;
    const unsigned char* ra = arec.empty() ? nullptr : arec.data() + t * AS;
    if (pos)
      for (int c = 0; c < 9; c++) pos[9 * t + c] = real_at(rp + c * rs);
    if (wfn)
      for (int c = 0; c < 3; c++) wfn[3 * t + c] = real_at(ra + c * rs);
    if (cell) std::memcpy(cell + t, ra + 3 * rs, 4);
    if (edges) std::memcpy(edges + t, fp64 ? ra + 3 * rs + 4 : rp + 9 * rs, 4);
  }
  if (vn) {
    // the mesh keeps no per-triangle vertex normals; they are gathered now
    // with the same sums the resolve pass uses for nearest triangles
    if (c3d::ensure_init(L, 0)) return 1;
    if (L.main.vntmp.ensure(n * 9 * rs) != hipSuccess) {
      c3d::log_error("mesh_copy: allocation failed");
      return 1;
    }
    if (c3d::smooth_object(L.main, *mesh, obj, L.main.vntmp.p)) return 1;
    std::vector<unsigned char> raw(n * 9 * rs);
    if (hipMemcpy(raw.data(), L.main.vntmp.p, n * 9 * rs, hipMemcpyDeviceToHost) != hipSuccess) {
      c3d::log_error("mesh_copy: device copy failed");
      return 1;
    }
    for (uint64_t q = 0; q < n * 9; q++) vn[q] = real_at(raw.data() + q * rs);
  }
  return 0;
}

int contour3d_amd_mesh_normals(const contour3d_amd_mesh_t* mesh, size_t obj, void* vn) {
  if (!mesh || !vn || obj >= mesh->count.size()) {
    c3d::log_error("mesh_normals: bad arguments");
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  const uint64_t n = mesh->count[obj];
  if (n == 0) return 0;
  const size_t bytes = size_t(n) * 9 * (mesh->precision == CONTOUR3D_AMD_FP64 ? 8 : 4);
  if (c3d::is_device_ptr(vn)) return c3d::smooth_object(L.main, *mesh, obj, vn);
  if (L.main.vntmp.ensure(bytes) != hipSuccess) {
    c3d::log_error("mesh_normals: allocation failed");
    return 1;
  }
  if (c3d::smooth_object(L.main, *mesh, obj, L.main.vntmp.p)) return 1;
  if (hipMemcpy(vn, L.main.vntmp.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    c3d::log_error("mesh_normals: device copy failed");
    return 1;
  }
  return 0;
}

int contour3d_amd_mesh_compare(const contour3d_amd_mesh_t* a, const contour3d_amd_mesh_t* b, size_t obj,
                               double stats[8]) {
  if (!a || !b || !stats || obj >= a->count.size() || obj >= b->count.size() ||
      a->precision != CONTOUR3D_AMD_FP64) {
    c3d::log_error("mesh_compare: bad arguments (a must be an FP64 mesh holding object %zu)", obj);
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  const uint64_t n = a->count[obj];
  stats[0] = double(n);
  for (int k = 1; k < 8; k++) stats[k] = 0.0;
  if (a->count[obj] != b->count[obj]) {
    stats[1] = double(n > b->count[obj] ? n - b->count[obj] : b->count[obj] - n);
    return 0;
  }
  c3d::Frame& F = L.main;
  if (F.vntmp.ensure(8 * sizeof(unsigned long long)) != hipSuccess) {
    c3d::log_error("mesh_compare: allocation failed");
    return 1;
  }
  unsigned long long* out = F.vntmp.as<unsigned long long>();
  hipError_t e = hipMemsetAsync(out, 0, 8 * sizeof(unsigned long long), F.st);
  const c3d::MeshView<double> va = c3d::view_of<double>(*a);
  const c3d::MeshView<double> vb64 = c3d::view_of<double>(*b);
  const c3d::MeshView<float> vb32 = c3d::view_of<float>(*b);
  const bool b32 = b->precision == CONTOUR3D_AMD_FP32;
  if (e == hipSuccess)
    e = c3d::launch_mesh_compare(va, a->meta.as<c3d::ObjMeta>() + obj, b32 ? static_cast<const void*>(&vb32)
                                                                            : static_cast<const void*>(&vb64),
                                 b32 ? 1 : 0, b->meta.as<c3d::ObjMeta>() + obj, n, out, F.st);
  unsigned long long h[8] = {};
  if (e == hipSuccess) e = hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, F.st);
  if (e == hipSuccess) e = hipStreamSynchronize(F.st);
  if (e != hipSuccess) {
    c3d::log_error("mesh_compare failed: %s", hipGetErrorString(e));
    return 1;
  }
  double verr, amax;
  std::memcpy(&verr, &h[1], sizeof(double));
  std::memcpy(&amax, &h[2], sizeof(double));
  stats[1] = double(h[0]);
  stats[2] = verr;
  stats[3] = amax;
  stats[4] = h[4] ? double(h[3]) * 1e-9 / double(h[4]) : 0.0;
  stats[5] = double(h[4]);
  stats[6] = double(h[5]);
  stats[7] = double(h[6]);
  return 0;
}

int contour3d_amd_mesh_render(const contour3d_amd_mesh_t* mesh, const contour3d_vector_t* camera_position,
                              const contour3d_vector_t* light_direction, const size_t screen_sizes[2],
                              const contour3d_vector_t* screen_center, const contour3d_vector_t screen_local[2],
                              const contour3d_color_t* bg_color, size_t num_lines,
                              const contour3d_line_obj_t* line_objs, uint8_t* rgb, double* depth) {
  if (!mesh || !rgb) {
    c3d::log_error("mesh_render: mesh and rgb must not be NULL");
    return 1;
  }
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (c3d::ensure_init(L, 0)) return 1;
  if (!c3d::args_ok(camera_position, light_direction, screen_sizes, screen_center, screen_local, bg_color))
    return 1;
  std::vector<c3d::Image> imgs{c3d::make_image(camera_position, light_direction, screen_sizes, screen_center,
                                               screen_local, bg_color, 0, nullptr, num_lines, line_objs, rgb,
                                               depth, false)};
  return c3d::render_mesh(L, L.main, *mesh, imgs);
}

int contour3d_amd_mesh_free(contour3d_amd_mesh_t* mesh) {
  if (!mesh) return 0;
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  if (L.ready) (void)hipStreamSynchronize(L.main.st);
  mesh->release();
  delete mesh;
  return 0;
}

int contour3d_amd_server_start(const contour3d_amd_server_opts_t* opts) {
  auto& L = lib();
  std::lock_guard<std::mutex> gs(c3d::g_server_mu);
  {
    std::lock_guard<std::mutex> g(L.mu);
    if (c3d::ensure_init(L, 0)) return 1;
    if (L.server) {
      c3d::log_error("server already running");
      return 1;
    }
  }
  auto* S = new c3d::Server();
  if (opts) S->opts = *opts;
  if (S->opts.max_images == 0) S->opts.max_images = 16;
  if (S->opts.hbm_budget == 0) S->opts.hbm_budget = uint64_t(200) * 1000 * 1000 * 1000;
  if (S->opts.max_wait_ms < 0) S->opts.max_wait_ms = 0;
  for (auto& F : S->frames) {
    const hipError_t e = F.create();
    if (e != hipSuccess) {
      c3d::log_error("server: stream creation failed: %s", hipGetErrorString(e));
      for (auto& G : S->frames) G.release();
      delete S;
      return 1;
    }
  }
  if (hipStreamCreateWithFlags(&S->copy_st, hipStreamNonBlocking) != hipSuccess) {
    c3d::log_error("server: stream creation failed");
    for (auto& G : S->frames) G.release();
    delete S;
    return 1;
  }
  {
    std::lock_guard<std::mutex> g(L.mu);
    L.server = S;
  }
  S->worker = std::thread(c3d::server_loop, std::ref(L), std::ref(*S));
  S->writer = std::thread(c3d::writer_loop, std::ref(*S));
  return 0;
}

namespace c3d {

// The checks run_frame makes of one image (frame_setup, tess_prepare,
// frame_lines), made at submit.
int validate_request(const contour3d_amd_image_t& im, bool ppm) {
  const size_t W = im.screen_sizes[0], H = im.screen_sizes[1];
  if (W == 0 || H == 0 || W > (size_t(1) << 16) || H > (size_t(1) << 16) ||
      uint64_t(W) * uint64_t(H) >= (uint64_t(1) << 31)) {
    log_error("server_submit: image size %zu x %zu outside 1..65536 per axis and 2^31 pixels", W, H);
    return 1;
  }
  if (!ppm && !im.rgb) {
    log_error("server_submit: rgb output pointer is NULL");
    return 1;
  }
  if (im.num_contours && !im.contour_objs) {
    log_error("server_submit: contour objects pointer is NULL");
    return 1;
  }
  for (size_t o = 0; o < im.num_contours; o++)
    if (validate_obj(im.contour_objs[o], o)) return 1;
  if (im.num_lines && !im.line_objs) {
    log_error("server_submit: line objects pointer is NULL");
    return 1;
  }
  for (size_t l = 0; l < im.num_lines; l++) {
    if (im.line_objs[l].nitems < 2 || !im.line_objs[l].converter) {
      log_error("line object %zu: needs at least two points and a converter", l);
      return 1;
    }
  }
  return 0;
}

// submit and submit_ppm: queues one request (fname: PPM output)
int server_submit(const contour3d_amd_image_t* image, int precision, const char* fname, uint64_t* ticket) {
  auto& L = lib();
  std::lock_guard<std::mutex> gs(g_server_mu);
  Server* S = L.server;
  if (!S || !image || !ticket) {
    log_error("server_submit: server not running or NULL argument");
    return 1;
  }
  if (precision != CONTOUR3D_AMD_FP64 && precision != CONTOUR3D_AMD_FP32) {
    log_error("unknown precision %d", precision);
    return 1;
  }
  // what run_frame would reject fails here, at submit, so that a frame in
  // flight never carries a request that cannot render
  if (validate_request(*image, fname != nullptr)) return 1;
  auto r = std::make_shared<Request>();
  r->img = *image;
  if (S->opts.snapshot_fields && image->num_contours) {
    // copy every distinct field to a device snapshot before returning, so
    // the caller may update its fields in place at once
    r->objs.assign(image->contour_objs, image->contour_objs + image->num_contours);
    std::vector<std::pair<const double*, double*>> done;
    int bad = 0;
    for (size_t o = 0; o < r->objs.size() && !bad; o++) {
      contour3d_contour_obj_t& ob = r->objs[o];
      double* dst = nullptr;
      for (const auto& m : done)
        if (m.first == ob.array) dst = m.second;
      if (!dst) {
        const size_t bytes = ob.glsizes[0] * ob.glsizes[1] * ob.glsizes[2] * sizeof(double);
        size_t got = 0;
        void* p = snap_take(*S, bytes, got);
        if (!p) {
          log_error("server_submit: no device snapshot of %zu bytes (allocation failed, or the snapshots "
                    "held would exceed the HBM budget)", bytes);
          bad = 1;
          break;
        }
        r->snaps.push_back({p, got});
        if (hipMemcpyAsync(p, ob.array, bytes, hipMemcpyDefault, S->copy_st) != hipSuccess) {
          log_error("server_submit: snapshot copy failed");
          bad = 1;
          break;
        }
        dst = static_cast<double*>(p);
        done.push_back({ob.array, dst});
      }
      ob.array = dst;
    }
    if (hipStreamSynchronize(S->copy_st) != hipSuccess && !bad) {
      log_error("server_submit: snapshot copy failed");
      bad = 1;
    }
    if (bad) {
      snap_give(*S, *r);
      return 1;
    }
    r->img.contour_objs = r->objs.data();
  }
  if (fname) {
    const size_t W = image->screen_sizes[0], H = image->screen_sizes[1];
    const size_t bytes = W * H * 3;
    r->fname = fname;
    r->pinned = pinned_take(*S, bytes, r->pinned_bytes);
    if (!r->pinned) {
      log_error("server_submit_ppm: cannot allocate a pinned image of %zu bytes", bytes);
      snap_give(*S, *r);
      return 1;
    }
    r->img.rgb = r->pinned;
  }
  r->precision = precision;
  r->bytes = estimate_bytes(*S, r->img, precision);
  r->t_submit = now_s();
  std::lock_guard<std::mutex> g(S->m);
  if (S->stop) {
    log_error("server_submit: server stopping");
    pinned_give(*S, r->pinned, r->pinned_bytes);
    snap_give(*S, *r);
    return 1;
  }
  r->ticket = S->next_ticket++;
  if (S->t_first == 0) S->t_first = r->t_submit;
  S->queue.push_back(r);
  S->tickets[r->ticket] = r;
  *ticket = r->ticket;
  S->cv_work.notify_all();
  return 0;
}

}  // namespace c3d

int contour3d_amd_server_submit(const contour3d_amd_image_t* image, int precision, uint64_t* ticket) {
  return c3d::server_submit(image, precision, nullptr, ticket);
}

int contour3d_amd_server_submit_ppm(const contour3d_amd_image_t* image, int precision, const char* fname,
                                    uint64_t* ticket) {
  if (!fname) {
    c3d::log_error("server_submit_ppm: NULL file name");
    return 1;
  }
  return c3d::server_submit(image, precision, fname, ticket);
}

int contour3d_amd_server_wait(uint64_t ticket, double* latency_ms) {
  auto& L = lib();
  c3d::Server* S;
  {
    std::lock_guard<std::mutex> gs(c3d::g_server_mu);
    S = L.server;
  }
  if (!S) {
    c3d::log_error("server_wait: server not running");
    return 1;
  }
  std::unique_lock<std::mutex> lk(S->m);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) {
    c3d::log_error("server_wait: unknown ticket %llu", (unsigned long long)ticket);
    return 1;
  }
  std::shared_ptr<c3d::Request> r = it->second;
  S->cv_done.wait(lk, [&] { return r->status >= 0; });
  S->tickets.erase(ticket);
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  return r->status;
}

int contour3d_amd_server_stats(contour3d_amd_server_stats_t* st) {
  auto& L = lib();
  std::lock_guard<std::mutex> gs(c3d::g_server_mu);
  c3d::Server* S = L.server;
  if (!S || !st) {
    c3d::log_error("server_stats: server not running or NULL argument");
    return 1;
  }
  std::lock_guard<std::mutex> g(S->m);
  std::vector<double> v = S->lat_ms;
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) {
    if (v.empty()) return 0.0;
    const size_t k = size_t(q * double(v.size() - 1) + 0.5);
    return v[std::min(k, v.size() - 1)];
  };
  st->completed = S->completed;
  st->failed = S->failed;
  st->batches = S->batches;
  st->mean_batch = S->batches ? double(S->images) / double(S->batches) : 0.0;
  st->p50_ms = pct(0.50);
  st->p99_ms = pct(0.99);
  st->max_ms = v.empty() ? 0.0 : v.back();
  st->frames_per_s = S->t_last > S->t_first ? double(S->completed) / (S->t_last - S->t_first) : 0.0;
  st->peak_admitted_bytes = S->peak;
  return 0;
}

int contour3d_amd_server_stop(void) {
  auto& L = lib();
  c3d::Server* S;
  {
    std::lock_guard<std::mutex> gs(c3d::g_server_mu);
    S = L.server;
    if (!S) return 0;
    std::lock_guard<std::mutex> g(S->m);
    S->stop = true;
    S->cv_work.notify_all();
  }
  S->worker.join();
  // every rendered PPM request is queued for the writer by now
  {
    std::lock_guard<std::mutex> g(S->m);
    S->writer_stop = true;
    S->cv_write.notify_all();
  }
  S->writer.join();
  {
    std::lock_guard<std::mutex> gs(c3d::g_server_mu);
    std::lock_guard<std::mutex> g(L.mu);
    L.server = nullptr;
  }
  for (auto& F : S->frames) F.release();
  for (auto& b : S->pinned_all) (void)hipHostFree(b.first);
  for (auto& b : S->snap_all) (void)hipFree(b.first);
  if (S->copy_st) (void)hipStreamDestroy(S->copy_st);
  delete S;
  return 0;
}

int contour3d_amd_profile_enable(int enable) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  L.main.prof.on = enable != 0;
  return 0;
}

int contour3d_amd_profile_read(double ms[CONTOUR3D_AMD_NSTAGES], uint64_t calls[CONTOUR3D_AMD_NSTAGES],
                               uint64_t* ntri_total) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  for (int s = 0; s < CONTOUR3D_AMD_NSTAGES; s++) {
    if (ms) ms[s] = L.main.prof.ms[s];
    if (calls) calls[s] = L.main.prof.calls[s];
  }
  if (ntri_total) *ntri_total = L.main.prof.ntri;
  return 0;
}

int contour3d_amd_profile_reset(void) {
  auto& L = lib();
  std::lock_guard<std::mutex> g(L.mu);
  for (int s = 0; s < CONTOUR3D_AMD_NSTAGES; s++) {
    L.main.prof.ms[s] = 0;
    L.main.prof.calls[s] = 0;
  }
  L.main.prof.ntri = 0;
  return 0;
}

}  // extern "C"
