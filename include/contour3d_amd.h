/*
 * contour3d_amd.h -- extensions of the MI355X Contour3D library.
 *
 * contour3d.h stays the drop-in interface.  This header adds what an
 * embedding solver or a serving process needs beyond it: device-resident
 * images, an FP32-geometry form, stage-level access to the triangle mesh,
 * profiling, and explicit device/memory control.  None of these entry
 * points exists in the reference; each says which reference behaviour it
 * exposes.
 *
 * Conventions shared with contour3d_execute: 0 = success, 1 = failure with
 * a "[CONTOUR3D ERROR] " line on stderr.  Pointers documented "host or
 * device" are classified with hipPointerGetAttributes; device pointers are
 * read or written in place.  Calls are serialised by a library mutex.
 */
#if !defined(CONTOUR3D_AMD_H)
#define CONTOUR3D_AMD_H

#include <stddef.h>
#include <stdint.h>
#include "contour3d.h"

#if defined(__cplusplus)
extern "C" {
#endif

/* Arithmetic forms. */
#define CONTOUR3D_AMD_FP64 0  /* every step in FP64: bit-identical to the reference */
#define CONTOUR3D_AMD_FP32 1  /* FP64 classification and edge tests, FP32 geometry,
                                 depth and shading */

/* Select the device and create the library's stream.  Optional: every
 * entry point initialises device 0 on first use. */
extern int contour3d_amd_init(int device);

/* Free every device buffer, cached geometry and the stream.  The library
 * keeps device buffers between calls (grow-only) so that repeated frames do
 * not allocate; this returns them. */
extern int contour3d_amd_finalize(void);

/* Arithmetic form used by contour3d_execute (default CONTOUR3D_AMD_FP64). */
extern int contour3d_amd_set_precision(int precision);

/* The identity coordinate map.  Passing it as `converter` skips the
 * one-time lattice check that every other converter gets (the reference
 * calls the converter per tetra vertex, triangulate.c:170; this library
 * evaluates it once per lattice point and caches the result per geometry). */
extern contour3d_vector_t contour3d_amd_identity(const contour3d_vector_t orthogonal);

/*
 * contour3d_execute with the image returned instead of written to a file.
 *   rgb    W*H*3 bytes in PPM order (top row first, output.c:113-123);
 *          host or device pointer
 *   depth  optional (may be NULL), W*H doubles, row j = screen y index j:
 *          the final z-buffer, -DBL_MAX where nothing was drawn; host or device
 *   precision  CONTOUR3D_AMD_FP64 or CONTOUR3D_AMD_FP32
 */
extern int contour3d_amd_execute_rgb(
    const contour3d_vector_t * camera_position,
    const contour3d_vector_t * light_direction,
    const size_t screen_sizes[2],
    const contour3d_vector_t * screen_center,
    const contour3d_vector_t screen_local[2],
    const contour3d_color_t * bg_color,
    size_t num_contours,
    const contour3d_contour_obj_t * contour_objs,
    size_t num_lines,
    const contour3d_line_obj_t * line_objs,
    uint8_t * rgb,
    double * depth,
    int precision
);

/*
 * contour3d_amd_execute_rgb without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc --
 * this variant calls no hipMalloc, hipFree or other allocating routine.  Its
 * intermediate buffers (tile counts and starts, job tables, the cell store,
 * z-buffer, fragment list, line fragments, lattice grids) are regions of one
 * static device array of 128 MiB that each GPU the library runs on holds
 * until the process ends; contour3d_amd_init (or the first call of any
 * other entry point) sets it up, and this call returns 1 before that.
 * The caller passes the inputs and outputs:
 *   fields     device pointers (read in place), with the
 *              contour3d_amd_identity converter
 *   rgb        device, W*H*3 bytes in PPM row order
 *   depth      device W*H doubles, or NULL
 * Limits (1 with an error line beyond them): 2^21 pixels (1920x1080 fits),
 * 2^24 triangles in FP32 (FP64 frames store nothing per triangle), 64
 * contour objects, 2^20 tessellation tiles (about 2^31 cells), 2^18 line
 * fragments, 65536 lattice coordinates; an FP64 frame with more than 2^19
 * covered pixel centres takes the slower z-test passes.  Results are
 * bit-identical to contour3d_amd_execute_rgb.  Every call on a GPU shares
 * that GPU's array and keeps the last call's contents; calls are serialised
 * by the library mutex, and calls that overlapped on one GPU would race.
 */
extern int contour3d_amd_execute_rgb_noalloc(
    const contour3d_vector_t * camera_position,
    const contour3d_vector_t * light_direction,
    const size_t screen_sizes[2],
    const contour3d_vector_t * screen_center,
    const contour3d_vector_t screen_local[2],
    const contour3d_color_t * bg_color,
    size_t num_contours,
    const contour3d_contour_obj_t * contour_objs,
    size_t num_lines,
    const contour3d_line_obj_t * line_objs,
    uint8_t * rgb,
    double * depth,
    int precision
);

/* ---- batches ------------------------------------------------------------ */

/* One image of a batch: the arguments of contour3d_amd_execute_rgb. */
typedef struct {
  contour3d_vector_t camera_position;
  contour3d_vector_t light_direction;
  size_t screen_sizes[2];
  contour3d_vector_t screen_center;
  contour3d_vector_t screen_local[2];
  contour3d_color_t bg_color;
  size_t num_contours;
  const contour3d_contour_obj_t * contour_objs;
  size_t num_lines;
  const contour3d_line_obj_t * line_objs;
  uint8_t * rgb;    /* W*H*3, host or device */
  double * depth;   /* W*H or NULL, host or device */
} contour3d_amd_image_t;

/*
 * Render several images as one frame: every contour object of every image
 * is tessellated by one pass of each tessellation kernel, and every image is
 * rasterised and resolved by one launch per stage.  Each image is identical
 * to what contour3d_amd_execute_rgb returns for it alone.  Images whose
 * contour lists are equal (same arrays, lattices, converters, thresholds,
 * colours, in order) share one tessellation: several cameras on one scene
 * cost one tessellation.  Limits: fewer than 2^31 pixels per batch, and in
 * FP32 2^32 - 2^24 triangles per batch (FP64 frames name triangles by
 * object, cell and slot and have no such limit).
 */
extern int contour3d_amd_execute_batch(size_t num_images, const contour3d_amd_image_t * images, int precision);

/* ---- serving ------------------------------------------------------------ */

typedef struct {
  size_t max_images;    /* images per batch at most (0: 16) */
  uint64_t hbm_budget;  /* estimated device bytes of the frames in flight at most, and separately the
                           bytes of field snapshots held (0: 200e9) */
  double max_wait_ms;   /* while the device is idle, a request waits this long for others (0: none) */
  int snapshot_fields;  /* 1: submit copies every field of the request (host or device) into a device
                           snapshot before it returns, so a solver may update its fields in place at
                           once; the request renders the snapshot.  0: fields are read in place. */
} contour3d_amd_server_opts_t;

typedef struct {
  uint64_t completed, failed, batches;
  double mean_batch;                /* images per batch */
  double p50_ms, p99_ms, max_ms;    /* from submit to the image in the caller's buffer */
  double frames_per_s;              /* completed images / (last completion - first submit) */
  uint64_t peak_admitted_bytes;
} contour3d_amd_server_stats_t;

/*
 * A request loop with continuous batching.  Submitted images queue in
 * arrival order; a worker thread batches them (same arithmetic form, up to
 * max_images, within the HBM budget) into frames rendered exactly as
 * contour3d_amd_execute_batch renders them, so each result is the same
 * alone and in any mix.  Two frames are in flight at once: requests that
 * arrive while one frame runs start on the other as soon as it is free.
 *
 * Every buffer an image points to (contour object array, fields, grids,
 * line objects, rgb, depth) must stay valid and unchanged from submit until
 * its wait returns; with snapshot_fields, the contour object array and the
 * fields only until submit returns.  Fields updated in place on the device
 * must be complete (the updating stream synchronised) before the submit
 * that reads them.  Submit checks each request as a frame would (sizes,
 * objects, lines) and returns 1 for one that could not render; a request
 * whose frame fails on the device fails alone.  start/stop must not run
 * concurrently with submit, wait or stats.
 */
extern int contour3d_amd_server_start(const contour3d_amd_server_opts_t * opts);
extern int contour3d_amd_server_submit(const contour3d_amd_image_t * image, int precision, uint64_t * ticket);
/* As _submit, but the image goes to `fname` as the binary PPM that
 * contour3d_execute writes (P6, output.c:125-140); image->rgb is ignored.
 * The server renders into a pinned host image of its own, and a writer
 * thread writes the file once the frame's copies are done, so the file I/O
 * of one frame overlaps the rendering of the next.  The wait returns once
 * the file is written; 1 if it could not be.  No counterpart in the
 * reference, whose contour3d_execute writes the file before it returns
 * (output.c:113-149): it is the asynchronous output path for a serving
 * process. */
extern int contour3d_amd_server_submit_ppm(const contour3d_amd_image_t * image, int precision, const char * fname,
                                           uint64_t * ticket);
/* Blocks until the request is done; 0 = rendered, 1 = failed.  Each ticket
 * is waited for once. */
extern int contour3d_amd_server_wait(uint64_t ticket, double * latency_ms);
extern int contour3d_amd_server_stats(contour3d_amd_server_stats_t * stats);
/* Renders everything still queued, then stops the worker. */
extern int contour3d_amd_server_stop(void);

/* ---- stage-level mesh access (tessellation + smoothing) --------------- */

typedef struct contour3d_amd_mesh contour3d_amd_mesh_t;

/* Tessellate and smooth every contour object (objects that share a field,
 * grids and converter share one pass over the field).  Triangles of object
 * n follow those of object n-1; within an object they are in the
 * reference's serial order (cell index k, j, i, then the slot order of
 * triangulate.c:325-338). */
extern int contour3d_amd_mesh_build(
    const contour3d_contour_obj_t * contour_objs,
    size_t num_contours,
    int precision,
    contour3d_amd_mesh_t ** mesh
);

/* Number of triangles of object `obj` (all cells, rendered or not). */
extern int contour3d_amd_mesh_count(const contour3d_amd_mesh_t * mesh, size_t obj, size_t * ntri);

/*
 * Copy object `obj`'s triangles to host arrays (any may be NULL), as FP64:
 *   pos   ntri*9: vertex v, axis d at [9t + 3v + d]
 *   wfn   ntri*3: area * unit face normal (triangulate.c:75-93, normal.c:39-40)
 *   vn    ntri*9: unit vertex normals; only triangles of interior cells
 *         (index 1..n-3 per axis, contour/main.c:92-93) have them, other
 *         rows are unspecified
 *   cell  ntri: linear cell index (k*ncy + j)*ncx + i
 *   edges ntri: three 5-bit cell-local lattice edge ids (triangulate.c:23-32)
 *         in bits 0-14, bit 15 set for triangles of interior cells
 */
extern int contour3d_amd_mesh_copy(const contour3d_amd_mesh_t * mesh, size_t obj, double * pos,
                                   double * wfn, double * vn, uint32_t * cell, uint32_t * edges);

/* The smoothing stage on its own (normal.c:46-165): the unit vertex normals
 * of every triangle of object `obj`, 9 values per triangle in the mesh's
 * precision (doubles for an FP64 mesh, floats for an FP32 one; vertex v,
 * axis d at [9t + 3v + d]), into `vn`, host or device.  Each is the sum of
 * area * unit face normal over the triangles of the cells sharing the
 * vertex's lattice edge that have a vertex on it, in serial order, from +0,
 * normalised; rows of triangles outside the rendered interior are NaN.
 * Timed as CONTOUR3D_AMD_STAGE_SMOOTH when profiling is on. */
extern int contour3d_amd_mesh_normals(const contour3d_amd_mesh_t * mesh, size_t obj, void * vn);

/* Render a built mesh (all its objects, in order) and then the lines, as
 * contour3d_amd_execute_rgb does. */
extern int contour3d_amd_mesh_render(
    const contour3d_amd_mesh_t * mesh,
    const contour3d_vector_t * camera_position,
    const contour3d_vector_t * light_direction,
    const size_t screen_sizes[2],
    const contour3d_vector_t * screen_center,
    const contour3d_vector_t screen_local[2],
    const contour3d_color_t * bg_color,
    size_t num_lines,
    const contour3d_line_obj_t * line_objs,
    uint8_t * rgb,
    double * depth
);

/* Compare object `obj` of an FP64 mesh `a` with the same object of mesh
 * `b` (built from the same contour objects, FP64 or FP32), on the device:
 *   stats[0] triangles of the object in a
 *   stats[1] topology mismatches (triangles whose cell or edge ids differ,
 *            or the difference of the triangle counts)
 *   stats[2] max |vertex coordinate difference|
 *   stats[3] max vertex-normal angle, degrees (interior triangles)
 *   stats[4] mean vertex-normal angle, degrees
 *   stats[5] vertex normals compared (non-finite ones skipped)
 *   stats[6] vertex normals off by more than 1 degree
 *   stats[7] vertex normals off by more than 10 degrees
 * The tool behind the FP32 form's stated errors. */
extern int contour3d_amd_mesh_compare(const contour3d_amd_mesh_t * a, const contour3d_amd_mesh_t * b, size_t obj,
                                      double stats[8]);

extern int contour3d_amd_mesh_free(contour3d_amd_mesh_t * mesh);

/* ---- tuning ------------------------------------------------------------- */

/* Launch-shape and capacity parameters.  None changes a result: every
 * setting renders the same bits. */
enum {
  CONTOUR3D_AMD_TUNE_TESS_GRID = 0,     /* max workgroups of the tessellation kernels (count, the fused
                                           tessellation + raster, emit); each walks several tiles
                                           (0: the defaults, 128 per CU, and one per tile for emit) */
  CONTOUR3D_AMD_TUNE_RASTER_GRID = 1,   /* max workgroups of the raster passes (0 = one per 256 triangles) */
  CONTOUR3D_AMD_TUNE_FRAG_CAPACITY = 2, /* FP64 fragment list entries, 24 B each (0: the default, 2^24; grows by
                                           itself after an overflow; an overflowing frame takes the slower
                                           z-test passes over the geometry) */
  CONTOUR3D_AMD_TUNE_GRAPHS = 3,        /* 1: contour3d_execute, _execute_rgb and _execute_batch capture
                                           their launch sequence as a HIP graph and replay it while the
                                           frame layout (image sizes, objects' fields and lattices,
                                           buffer capacities) repeats; thresholds, colours, cameras and
                                           lines may change between replays.  Default 0. */
  CONTOUR3D_AMD_TUNE_FAIL_ALLOC = 4,    /* failure injection for testing the error contract: n > 0 makes
                                           the n-th device or pinned allocation from now fail as an
                                           out-of-memory would (once); the call that needed it returns 1
                                           with the error line, and the library stays usable.  0 disarms. */
  CONTOUR3D_AMD_TUNE_TESS_ROWS = 5,     /* tessellation brick: rows of cells per tile, at most 2048 cells
                                           (0: as many rows as fit) */
  CONTOUR3D_AMD_TUNE_TESS_ZPLANES = 6   /* tessellation brick: z-planes one workgroup marches through per
                                           work item (0: the default, 4) */
};

extern int contour3d_amd_tune(int param, int64_t value);

/* ---- profiling --------------------------------------------------------- */

/* Stages timed with HIP events on the library stream when profiling is on.
 * A frame needs the smoothed vertex normals (normal.c) only of the triangle
 * nearest to each pixel; they are gathered inside the RESOLVE stage, so the
 * SMOOTH stage stays at zero for frames. */
enum {
  CONTOUR3D_AMD_STAGE_COUNT = 0,   /* classify + per-tile triangle counts */
  CONTOUR3D_AMD_STAGE_SCAN = 1,    /* tile scan + totals */
  CONTOUR3D_AMD_STAGE_EMIT = 2,    /* triangle emission */
  CONTOUR3D_AMD_STAGE_SMOOTH = 3,  /* vertex normals of every triangle (contour3d_amd_mesh_normals; not used
                                      by frames) */
  CONTOUR3D_AMD_STAGE_RASTER = 4,  /* z-buffer passes incl. lines */
  CONTOUR3D_AMD_STAGE_RESOLVE = 5, /* vertex normals of the nearest triangles, colour, y-flip pack */
  CONTOUR3D_AMD_STAGE_FUSED = 6,   /* within RASTER of a frame: the fused tessellation + raster kernel alone */
  CONTOUR3D_AMD_NSTAGES = 7
};

extern int contour3d_amd_profile_enable(int enable);

/* Accumulated milliseconds and launch counts per stage since the last
 * reset; ntri_total receives the triangles emitted since the last reset. */
extern int contour3d_amd_profile_read(double ms[CONTOUR3D_AMD_NSTAGES],
                                      uint64_t calls[CONTOUR3D_AMD_NSTAGES],
                                      uint64_t * ntri_total);

extern int contour3d_amd_profile_reset(void);

#if defined(__cplusplus)
}
#endif

#endif /* CONTOUR3D_AMD_H */
