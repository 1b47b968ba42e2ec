/*
 * Single-rank decomposition types for building and testing this library on
 * its own.  An embedding solver compiles against its own SimpleDecomp
 * sdecomp.h instead; the library only needs the two type names that
 * contour3d.h mentions (the reference header includes sdecomp.h for the
 * same reason, include/contour3d.h:6) and never calls a SimpleDecomp
 * function.
 *
 * sdecomp_pencil_t is an int-sized enum, so contour3d_contour_obj_t has the
 * same layout whichever header provides it.  The info object is opaque.
 */
#if !defined(SDECOMP_H)
#define SDECOMP_H

typedef struct sdecomp_info_t_ sdecomp_info_t;

typedef enum {
  SDECOMP_XDIR = 0,
  SDECOMP_YDIR = 1,
  SDECOMP_ZDIR = 2
} sdecomp_dir_t;

typedef enum {
  SDECOMP_X1PENCIL = 0,
  SDECOMP_Y1PENCIL = 1,
  SDECOMP_Z1PENCIL = 2,
  SDECOMP_X2PENCIL = 3,
  SDECOMP_Y2PENCIL = 4,
  SDECOMP_Z2PENCIL = 5
} sdecomp_pencil_t;

#endif /* SDECOMP_H */
