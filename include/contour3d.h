/*
 * contour3d.h -- drop-in public C API of the MI355X Contour3D library.
 *
 * The prototype of contour3d_execute and the layouts of the four public
 * structs are kept identical to the reference interface
 * (include/contour3d.h:8-86), so an embedding solver that was compiled
 * against the reference header re-links against libcontour3d.so without
 * source changes.  Field order and member types below follow
 * include/contour3d.h:10-56 member for member; the names are part of the
 * ABI and therefore unchanged.
 *
 * The decomposition types (sdecomp_info_t, sdecomp_pencil_t) come from the
 * caller's SimpleDecomp header, exactly as in the reference
 * (include/contour3d.h:6).  This library runs the whole field on one GPU:
 * it never calls into SimpleDecomp, treats every contour object as the full
 * global array (local size == glsizes, offset == 0), and accepts but does
 * not dereference the sdecomp_info pointer.
 */
#if !defined(CONTOUR3D_H)
#define CONTOUR3D_H

#include <stddef.h> /* size_t */
#include <stdint.h> /* uint8_t */
#include "sdecomp.h"

#define CONTOUR3D_NDIMS 3

#if defined(__cplusplus)
extern "C" {
#endif

/* Cartesian (or user-orthogonal) 3-vector, include/contour3d.h:10-14. */
typedef struct {
  double x;
  double y;
  double z;
} contour3d_vector_t;

/* 8-bit RGB colour, include/contour3d.h:16-20. */
typedef struct {
  uint8_t r;
  uint8_t g;
  uint8_t b;
} contour3d_color_t;

/*
 * One iso-surface request, include/contour3d.h:22-40.
 *   pencil     decomposition pencil of `array` (ignored on one GPU)
 *   glsizes    number of lattice points in x, y, z
 *   grids      lattice coordinates per axis, lengths glsizes[0..2]
 *   converter  maps (grids[0][i], grids[1][j], grids[2][k]) to Cartesian
 *   threshold  iso-value; a corner is "above" when threshold < value
 *   color      surface colour
 *   array      FP64 field, x fastest: array[(k*ny + j)*nx + i]; host or
 *              device (HBM) pointer -- device pointers are used in place
 */
typedef struct {
  sdecomp_pencil_t pencil;
  size_t glsizes[CONTOUR3D_NDIMS];
  double * grids[CONTOUR3D_NDIMS];
  contour3d_vector_t (* converter) (
      const contour3d_vector_t orthogonal
  );
  double threshold;
  contour3d_color_t color;
  double * array;
} contour3d_contour_obj_t;

/*
 * One poly-line request, include/contour3d.h:42-56: the segment
 * edges[0] -> edges[1] is cut into nitems-1 pieces, each mapped through
 * converter and drawn with the given pixel width.
 */
typedef struct {
  contour3d_vector_t edges[2];
  size_t nitems;
  contour3d_vector_t (* converter) (
      const contour3d_vector_t orthogonal
  );
  contour3d_color_t color;
  double width;
} contour3d_line_obj_t;

/*
 * Render every contour object (in order) and then every line object into
 * one z-buffered image and write it as a binary PPM to `fname`.
 * Same prototype and contract as include/contour3d.h:58-86: returns 0 on
 * success, 1 on failure with a "[CONTOUR3D ERROR] " message on stderr; no
 * memory the call allocated outlives a failed call.
 */
extern int contour3d_execute (
    const sdecomp_info_t * const sdecomp_info,
    const contour3d_vector_t * const camera_position,
    const contour3d_vector_t * const camera_look_at,
    const contour3d_vector_t * const light_direction,
    const size_t screen_sizes[2],
    const contour3d_vector_t * const screen_center,
    const contour3d_vector_t screen_local[2],
    const contour3d_color_t * const bg_color,
    const size_t num_contours,
    const contour3d_contour_obj_t * const contour3d_contour_objs,
    const size_t num_lines,
    const contour3d_line_obj_t * const contour3d_line_objs,
    const char fname[]
);

#if defined(__cplusplus)
}
#endif

#endif /* CONTOUR3D_H */
